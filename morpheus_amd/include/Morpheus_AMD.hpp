// Morpheus_AMD.hpp -- umbrella header of the MI355X drop-in for the Morpheus
// SpMV path (the counterpart of core/src/Morpheus.hpp): spaces, containers,
// copy / mirrors, multiply, convert, the DynamicMatrix switch, the matrix
// analytics and MatrixMarket I/O, the device work served
// by the HIP kernels of libmorpheus_amd.so through include/morpheus_amd.h.
//
// Build: any C++17 compiler with the HIP runtime headers, e.g.
//   g++ -std=c++17 -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include
//       -I<repo>/include -I<repo>/morpheus_amd/include app.cpp
//       -L<repo>/morpheus_amd -lmorpheus_amd -L/opt/rocm/lib -lamdhip64
#pragma once

#include "Morpheus_AMD_Spaces.hpp"
#include "Morpheus_AMD_Containers.hpp"
#include "Morpheus_AMD_Algorithms.hpp"
#include "Morpheus_AMD_MatrixMarket.hpp"
