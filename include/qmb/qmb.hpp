// C++ facade of the MI355X ED engine: the reference's lattice-model, symmetry,
// quantum-operator and ED-driver class API (header-only, over the C ABI in
// include/qmb.h). Link with -lqmb.
#pragma once

#include "../qmb.h"
#include "ed.hpp"
#include "lattice.hpp"
#include "observables.hpp"
#include "operators.hpp"
#include "symmetry.hpp"
