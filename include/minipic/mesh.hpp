// Mesh (mesh.hpp) of the reference's class interface, so reference-style
// `#include "mesh.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_MESH_HPP_
#define MINIPIC_MESH_HPP_
#include "minipic.hpp"
#endif
