// kites.cuh -- include name of the reference's umbrella header
// (kites/kites.cuh), kept so reference user code compiles unchanged.
#include "kites.hpp"
