// PIC (PIC.hpp) of the reference's class interface, so reference-style
// `#include "PIC.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_PIC_HPP_
#define MINIPIC_PIC_HPP_
#include "minipic.hpp"
#endif
