// ParticleFill (particle_fill.hpp) of the reference's class interface, so reference-style
// `#include "particle_fill.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_PARTICLE_FILL_HPP_
#define MINIPIC_PARTICLE_FILL_HPP_
#include "minipic.hpp"
#endif
