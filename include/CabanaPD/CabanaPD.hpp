// CabanaPD.hpp -- umbrella header of the drop-in API (reference: src/CabanaPD.hpp).
// Link with libpdk.so; the compute path is the MI355X engine behind pdk.h.
#ifndef CABANAPD_DROPIN_HPP
#define CABANAPD_DROPIN_HPP

#include "CabanaPD_Boundary.hpp"
#include "CabanaPD_ForceModels.hpp"
#include "CabanaPD_Input.hpp"
#include "CabanaPD_Particles.hpp"
#include "CabanaPD_Solver.hpp"
#include "CabanaPD_Types.hpp"

#endif
