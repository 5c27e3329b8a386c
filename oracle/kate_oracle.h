/*
 * kate_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Kate's CPU simulate path, used as the checker for
 * the HIP library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; nothing in kate_amd/ links or calls it.
 *
 * What it restates (reference file:line):
 *   orc_apply_gate    CPUcode/GateComputing.hpp:6-128   (computeGate)
 *   orc_merge         CPUcode/GateMerger.hpp:8-55       (CPUmergeGate)
 *   orc_run           CPUcode/simulator.hpp:71-99,198-291 (execute + executeCommand,
 *                     14-qubit group staging, per-group thread fan-out)
 *   orc_init_basis    CPUcode/simulator.hpp:133-138
 *   orc_init_product  CPUcode/simulator.hpp:139-162
 *   orc_measure       CPUcode/simulator.hpp:163-197     (proba_state)
 *   orc_swap_bits     HIPcode/simulator.hpp:639-661,751-759 (a swap instruction
 *                     exchanges two qubit positions; Kate's CPU path ignores
 *                     swaps, its multi-GPU path performs them)
 *   orc_expect_z      no counterpart in Kate: <Z_S> = sum |a_x|^2 (-1)^popcount(x&S)
 *
 * Amplitudes are complex128, interleaved (re, im), little-endian qubit order:
 * bit q of the index is the value of (virtual) qubit q.
 */
#ifndef KATE_ORACLE_H
#define KATE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DENSE = 0, ORC_H = 2, ORC_CNOT = 3, ORC_CRK = 4, ORC_TOFFOLI = 5 };

/* Gate g: id[g], qubits[qoff[g] .. qoff[g]+nq[g]), optarg[g]; a dense gate's
 * 2^nq x 2^nq row-major complex matrix starts at mat[moff[g]] (in doubles). */
typedef struct {
    int32_t ngates;
    const int32_t* id;
    const int32_t* nq;
    const int32_t* qoff;
    const int32_t* qubits;
    const int32_t* optarg;
    const int64_t* moff;
    const double* mat;
} orc_gates;

/* Compiled program: groups[g] = (end-exclusive gate index, qubit bitmask);
 * instruction i is a compute (kind 1) or a swap list (kind 0) whose position
 * pairs are pairs[poff[i] .. poff[i]+2*npairs[i]). */
typedef struct {
    int32_t nqubits;
    orc_gates gates;
    int32_t ngroups;
    const int32_t* group_end;
    const uint64_t* group_mask;
    int32_t ninstr;
    const int32_t* instr_kind;
    const int32_t* instr_npairs;
    const int32_t* instr_poff;
    const int32_t* pairs;
} orc_program;

/* Returns 0 on success, <0 on a malformed gate. */
int orc_apply_gate(const orc_gates* gs, int g, int nlocal, double* local, const int32_t* bit_to_local);
int orc_merge(const orc_gates* gs, int first, int last, int32_t* out_qubits, double* out_mat);
void orc_init_basis(int nq, double* state);
void orc_init_product(int nq, const double* theta_phi, const int32_t* initial_permutation, double* state);
int orc_run(const orc_program* p, double* state, int nthreads);
void orc_swap_bits(int nq, double* state, int p1, int p2);
void orc_measure(int nq, const double* state, const int32_t* final_inverse_permutation, double* theta_phi);
void orc_expect_z(int nq, const double* state, int m, const uint64_t* masks, double* out);

#ifdef __cplusplus
}
#endif
#endif
