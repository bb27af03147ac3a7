// kites.hpp -- umbrella header of the MI355X vertex-centric engine
// (the counterpart of the reference's kites/kites.cuh).
#ifndef KITES_KITES_HPP_
#define KITES_KITES_HPP_

#include "common.hpp"
#include "device.hpp"
#include "memory.hpp"
#include "io.hpp"
#include "graph.hpp"
#include "graph_device.hpp"
#include "engine.hpp"
#include "batch.hpp"
#include "sanity_check.hpp"

#endif  // KITES_KITES_HPP_
