// Internal declarations shared by the translation units of libqfe.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/qfe.h"

namespace qfe {

// ---- errors ------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define QFE_HIP(call)                                        \
    do {                                                     \
        hipError_t qfe_e_ = (call);                          \
        if (qfe_e_ != hipSuccess) return ::qfe::hip_fail(qfe_e_, #call); \
    } while (0)

// ---- library context -----------------------------------------------------------
struct Tuning {
    bool tile_auto = true; // tile shape chosen by state size (tile_shape); false: the three values below
    int tile_qubits = 11;  // k: qubits held by one fused tile (2^k amplitudes in LDS)
    int low_qubits = 4;    // L: lowest qubits always in the tile (contiguous 2^L runs)
    int max_ops = 24;      // ops per fused pass before LDS traffic dominates
    int orbit_dims = 7;    // max dimensions of a register orbit (5, 6, 7: spread over 2, 4 or 8 lanes)
    uint64_t sector_words = 1ull << 25;  // max pair words of one sector template (256 MB)
};

// Pass shape for a state of n qubits: the explicit tuning, or by default the
// shape measured fastest for the size on MI355X.  A small state needs small
// tiles so that a pass has enough workgroups and each op a short per-thread
// chain: (8, 4, 64) up to 14 qubits (500 random gates at 12 qubits: 0.59 vs
// 0.93 ms), (10, 4, 32) at 15-20 (the 16-qubit UCCSD energy, 2688
// exponentials: 2.73 ms, against 3.38 ms with 8-qubit tiles), (11, 4, 24)
// from 21 on.  Register-orbit passes keep 24 ops: longer ones are bound by
// their arithmetic while the state sits in cache (the same UCCSD energy with
// 64: 4.85 ms).
struct TileShape {
    int k, low, max_ops, orbit_ops;
};
inline TileShape tile_shape(const Tuning& tu, int n) {
    if (!tu.tile_auto) return {tu.tile_qubits, tu.low_qubits, tu.max_ops, tu.max_ops};
    if (n <= 14) return {8, 4, 64, 24};
    if (n <= 20) return {10, 4, 32, 24};
    return {11, 4, 24, 24};
}

struct Counters {
    uint64_t fused = 0, single = 0, reduce = 0;
    // fused passes by kind, and read-write or expectation passes launched with
    // non-temporal accesses
    uint64_t tile = 0, orbit = 0, lane_orbit = 0, nt = 0;
    double bytes = 0.0;
};

struct Context {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Tuning tuning;
    Counters counters;
    // reusable device scratch for reductions and op descriptors
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // pinned host staging for small results
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
};

Context& ctx();
int ensure_init();
int scratch_reserve(size_t bytes, void** out);
int pinned_reserve(size_t bytes, void** out);
// Double-buffered pinned staging: fill *host (bytes), then commit copies it to
// dev on the library stream without blocking the host.
int staging_begin(size_t bytes, void** host);
int staging_commit(void* dev, size_t bytes);

}  // namespace qfe

struct qfe_state_s {
    int n = 0;
    uint64_t N = 0;
    double2* d = nullptr;
};

namespace qfe {

// ---- device op descriptors -------------------------------------------------------
// One operation as the kernels see it.  Every op acts on a set of "mixing"
// qubits (where it permutes or mixes amplitudes); everything else it needs
// (controls, diagonal phases, Z masks) is a per-index condition or sign.
enum OpKind : uint32_t {
    K_U1 = 0,     // 2x2 unitary on bit t, applied where (j & cmask) == cval
    K_DIAG1 = 1,  // diagonal 2x2 (d0 on bit t = 0, d1 on bit t = 1) where (j & cmask) == cval
    K_U2 = 2,     // general 4x4 on bits (t, c), basis 2*c_bit + t_bit
    K_PEXP = 3,   // exp(i angle P): cos(angle) I + i sin(angle) P, P = phase (-1)^{j.z} |j ^ x>
};

struct DevOp {
    uint32_t kind;
    uint32_t t;       // target bit (global) or local bit inside a fused tile
    uint32_t c;       // second bit for K_U2
    uint32_t hb;      // K_PEXP: highest set bit of x (global or local); K_U1 / K_U2: 1 when the matrix
                      // is an X / a qubit swap (a permutation a tile's index map absorbs)
    uint64_t x;       // K_PEXP: flip mask
    uint64_t z;       // K_PEXP: sign mask (in-tile part when fused)
    uint64_t cmask;   // condition mask (in-tile part when fused)
    uint64_t cval;    // condition value
    uint64_t omask;   // fused only: out-of-tile condition mask (tested on tile base)
    uint64_t oval;    // fused only: out-of-tile condition value
    uint64_t ozmask;  // fused only: out-of-tile z mask (sign from tile base)
    uint32_t tin;     // fused K_DIAG1: 1 if the diagonal bit lies inside the tile
    uint32_t pad;
    double2 m[16];    // K_U1: m[0..3] = m00 m01 m10 m11; K_DIAG1: m[0], m[1];
                      // K_U2: row-major 4x4; K_PEXP: m[0] = (cos, sin), m[1] = phase
};

// Host-side representation of an op before planning.
struct HostOp {
    DevOp op;            // global-bit form
    uint64_t mix;        // mixing qubits
    uint64_t support;    // every qubit the op reads (for commutation)
    int32_t src = -1;    // K_PEXP with a device-side angle: index into the caller's angle table
};

// Build HostOps from gates / op stream. Returns false + message on bad input.
int lower_gate(const qfe_gate& g, int n, std::vector<HostOp>& out);
int lower_pexp(const qfe_pauli& p, double angle, int n, std::vector<HostOp>& out);

// Execute a lowered op list on a state (plans fused passes and launches).
int run_ops(qfe_state_s* s, std::vector<HostOp>& ops);

// A planned op list kept for replay: the descriptor image (uploaded by the
// caller to device memory it owns) and the launch sequence over it.  With
// device_angles every K_PEXP lands in a fused, orbit or lane-orbit pass (never
// a standalone pass, whose angle is a kernel argument), and `angles` lists the
// byte offset in the image of each such op's (cos, sin) pair with the op's
// `src` tag, so a kernel can rewrite the angles before each replay.
struct OpPlan;
struct AngleRef {
    uint64_t off;
    int32_t src;
    uint32_t pad;
};
int plan_build(std::vector<HostOp>& ops, int n, bool device_angles, OpPlan** out, std::vector<char>& image,
               std::vector<AngleRef>& angles);
int plan_launch(const OpPlan* p, qfe_state_s* s, char* dev_image);
void plan_free(OpPlan* p);

// ---- reductions / Pauli sums ---------------------------------------------------
struct PauliGroup {
    uint64_t x;
    uint32_t first;  // index into term arrays
    uint32_t count;
    uint32_t hb;     // highest bit of x (0 if x == 0)
    uint32_t pad;
};

// Group terms by x mask (stable: first-appearance order), folding phase*coeff.
void group_terms(const qfe_pauli* terms, uint64_t count, std::vector<PauliGroup>& groups,
                 std::vector<uint64_t>& zs, std::vector<double2>& cs);

int expval_groups(const double2* bra, const double2* ket, int n, const std::vector<PauliGroup>& groups,
                  const std::vector<uint64_t>& zs, const std::vector<double2>& cs, double2* out_per_group);

uint32_t expval_chunks(int n, size_t ngroups);
// Asynchronous forms on device-resident group tables: per-group sums land in dout[0..ngroups).
int launch_expval_async(const double2* bra, const double2* ket, int n, const PauliGroup* dgroups, uint32_t ngroups,
                        const uint64_t* dzs, const double2* dcs, double2* partial, uint32_t nchunk, double2* dout);
int launch_apply_sum_async(const double2* in, double2* out, int n, const PauliGroup* dgroups, uint32_t ngroups,
                           const uint64_t* dzs, const double2* dcs);

// program introspection for the serving engine
int program_nqubit(qfe_program p);
uint64_t program_nparam(qfe_program p);
struct ProgramView {
    int n;
    uint64_t ref;
    uint64_t nparam;
    const std::vector<qfe_pauli>* gen_terms;
    const std::vector<uint64_t>* gen_off;
    const std::vector<int32_t>* gparam;
    const std::vector<double>* gscale;
    const std::vector<qfe_pauli>* ham_terms;
};
ProgramView program_view(qfe_program p);

// Packing of X-mask groups into orbit passes (span of the masks <= 4 dims).
struct EvPass {
    std::vector<uint32_t> groups;  // indices into the group list
    std::vector<uint64_t> basis;   // reduced basis, pivot = highest bit
};
void plan_expval(const std::vector<PauliGroup>& groups, int maxdim, std::vector<EvPass>& passes);

// device address of the fixed expectation scratch on `device` (current)
int noalloc_scratch_setup(int device);

// device helpers exposed across translation units
int launch_zero(double2* d, uint64_t N, hipStream_t st);
int launch_basis(double2* d, uint64_t N, uint64_t idx, hipStream_t st);
int launch_apply_pauli_sum(const double2* in, double2* out, int n, const std::vector<PauliGroup>& groups,
                           const std::vector<uint64_t>& zs, const std::vector<double2>& cs);

}  // namespace qfe
