// Compile-only stand-in for FHDeX's TurbForcingComp class (src_compressible_stag/TurbForcingComp.H).
#pragma once

class TurbForcingComp {};
