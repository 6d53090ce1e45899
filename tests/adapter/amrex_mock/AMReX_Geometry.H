// Compile-only stand-in for amrex::Geometry (see AMReX_MultiFab.H here).
#pragma once

#include "AMReX_MultiFab.H"

namespace amrex {

class Geometry {
public:
    Real CellSize(int) const { return 0.; }
    Real ProbLo(int) const { return 0.; }
    Box Domain() const { return Box{}; }
    Periodicity periodicity() const { return Periodicity{}; }
};

const Geometry& DefaultGeometry();

}  // namespace amrex
