// FLOAT, LO, GO and VECTOR (types.hpp) of the reference's class interface, so reference-style
// `#include "types.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_TYPES_HPP_
#define MINIPIC_TYPES_HPP_
#include "minipic.hpp"
#endif
