// ES (ES.hpp) of the reference's class interface, so reference-style
// `#include "ES.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_ES_HPP_
#define MINIPIC_ES_HPP_
#include "minipic.hpp"
#endif
