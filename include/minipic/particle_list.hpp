// ParticleList and Particle (particle_list.hpp, particle.hpp) of the reference's class interface, so reference-style
// `#include "particle_list.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_PARTICLE_LIST_HPP_
#define MINIPIC_PARTICLE_LIST_HPP_
#include "minipic.hpp"
#endif
