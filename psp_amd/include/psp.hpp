// psp.hpp -- C++ mirror of PSP's umbrella header (include/psp.hpp:13-16) on
// one MI355X.  Header-only; every call goes through the C ABI of psp_hip.h.
#pragma once

#include "psp/environment.hpp"
#include "psp/discretization.hpp"
#include "psp/dist_uniform_grid.hpp"
#include "psp/dist_helmholtz.hpp"
