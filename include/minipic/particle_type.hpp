// ParticleType and ParticleTypeList (particle_type.hpp) of the reference's class interface, so reference-style
// `#include "particle_type.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_PARTICLE_TYPE_HPP_
#define MINIPIC_PARTICLE_TYPE_HPP_
#include "minipic.hpp"
#endif
