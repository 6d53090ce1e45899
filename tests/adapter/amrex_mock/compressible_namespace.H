// Compile-only stand-in for FHDeX's compressible namespace (src_compressible/compressible_namespace.H).
#pragma once

#include "AMReX_MultiFab.H"

namespace compressible {
extern int transport_type;
extern amrex::Real zeta_ratio;
}  // namespace compressible
