// Compile-only stand-in for the part of AMReX's API that include/fhdx_amrex*.H use
// (declarations with AMReX's names and argument types; no behaviour).  It lets the CPU tests
// check that the adapters are valid C++ and call the C ABI of fhdx.h with the right arguments.
#pragma once

#include <array>
#include <string>
#include <vector>

#define AMREX_SPACEDIM 3
#define AMREX_GPU_MANAGED
#ifndef MAX_SPECIES
#define MAX_SPECIES 8
#endif
#ifndef MAX_REACTION
#define MAX_REACTION 8
#endif

namespace amrex {

using Real = double;

template <class T, int N>
struct GpuArray {
    T arr[N];
    T& operator[](int i) { return arr[i]; }
    const T& operator[](int i) const { return arr[i]; }
};

template <class T>
using Vector = std::vector<T>;

struct Dim3 {
    int x, y, z;
};

template <class T>
struct Array4 {
    T* p;
    Dim3 begin, end;
    int ncomp;
};

struct IntVect {
    int v[3];
    int operator[](int i) const { return v[i]; }
};

struct Box {};

struct BoxArray {
    BoxArray() = default;
    explicit BoxArray(const Box&) {}
    long size() const { return 1; }
};

struct DistributionMapping {
    DistributionMapping() = default;
    DistributionMapping(const BoxArray&, int) {}
};

struct Periodicity {};

class MFIter;

class MultiFab {
public:
    MultiFab() = default;
    MultiFab(const BoxArray&, const DistributionMapping&, int, int) {}
    void define(const BoxArray&, const DistributionMapping&, int, int) {}
    Array4<Real> array(const MFIter&) { return Array4<Real>{}; }
    int nComp() const { return 0; }
    int nGrow() const { return 0; }
    const BoxArray& boxArray() const { return ba_; }
    const DistributionMapping& DistributionMap() const { return dm_; }
    void FillBoundary(const Periodicity&) {}

private:
    BoxArray ba_;
    DistributionMapping dm_;
};

class MFIter {
public:
    explicit MFIter(const MultiFab&) {}
    bool isValid() const { return false; }
    MFIter& operator++() { return *this; }
};

[[noreturn]] void Abort(const std::string& msg);

namespace Gpu {
void* gpuStream();
}

namespace ParallelDescriptor {
int IOProcessorNumber();
int NProcs();
template <class T>
void Bcast(T*, std::size_t, int)
{
}
}  // namespace ParallelDescriptor

}  // namespace amrex
