// DataWarehouse (data_warehouse.hpp) of the reference's class interface, so reference-style
// `#include "data_warehouse.hpp"` works with -Iinclude/minipic.  All classes live in minipic.hpp.
#ifndef MINIPIC_DATA_WAREHOUSE_HPP_
#define MINIPIC_DATA_WAREHOUSE_HPP_
#include "minipic.hpp"
#endif
