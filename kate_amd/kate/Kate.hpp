// Kate.hpp -- drop-in umbrella header of the MI355X build.
//
// Replaces the reference's single-node umbrella (Kate.hpp:1-64): the same
// namespace, types, passes and Simulator, backed by libkate_hip.so.  Build a
// user program with
//   g++ -std=c++17 -I<repo>/include -I<repo>/kate_amd/kate main.cpp \
//       -L<repo>/kate_amd -lkate_hip -Wl,-rpath,<repo>/kate_amd
#ifndef KATE_AMD_KATE_HPP
#define KATE_AMD_KATE_HPP

#define INITIALIZER
#define FINALIZER

#include "kate_types.hpp"
#include "kate_graph.hpp"
#include "kate_gatemath.hpp"
#include "kate_circuit.hpp"
#include "kate_simulator.hpp"
#include "kate_bench.hpp"

#endif
