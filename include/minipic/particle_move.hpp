// ParticleMove (particle_move.hpp) of the reference's class interface, so reference-style
// `#include "particle_move.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_PARTICLE_MOVE_HPP_
#define MINIPIC_PARTICLE_MOVE_HPP_
#include "minipic.hpp"
#endif
