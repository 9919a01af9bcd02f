// Internal declarations shared by the host runtime and the HIP kernels.
#ifndef LDSF_INTERNAL_H_
#define LDSF_INTERNAL_H_

#include <mutex>
#include <string>
#include <vector>

namespace ldsf {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string &msg);
void clear_error();

// The number of coordinate rows the caller's arrays hold, for the host-pointer
// calls this thread makes while the guard lives: a neighbour or centre index
// at or past it is refused (LDSF_ERR_ARG) before any species or coordinate
// row is read.  Without a guard the calls read max(index) + 1 rows, as the
// reference does.
long host_rows_limit();
struct HostRowsLimit {
    explicit HostRowsLimit(long rows);
    ~HostRowsLimit();
    long saved;
};

// Kinds, as SymmetryFunctions::add_descriptor numbers them
// (src/SymmetryFunctions.cpp:47-51).
enum Kind { G1 = 1, G2 = 2, G3 = 3, G4 = 4, G5 = 5 };

inline bool two_body(int k) { return k == G1 || k == G2 || k == G3; }
inline bool three_body(int k) { return k == G4 || k == G5; }

// Largest integer zeta taken by the exponentiation-by-squaring path.
constexpr int kMaxIntZeta = 1 << 20;

// Device view of a parameter set, passed by value to every kernel.  All
// pointers are device pointers into one allocation owned by the ldsf_sf.
// Rows are the reference's parameter rows in slot order; a row's output slot
// is slot + species_code * stride (two-body) or slot + pair_index * stride
// (three-body), stride being 0 unless the layout is species resolved.
struct SFTab {
    int n_species;
    int width;
    int n2;  // two-body rows
    int n3;  // three-body rows
    int species_resolved;
    int pad_;
    const double *rc;    // n_species * n_species, row-major
    const int *k2;       // kind per two-body row
    const double *a2;    // eta (g2) or kappa (g3)
    const double *b2;    // Rs (g2)
    const int *slot2;
    const int *str2;
    const int *k3;       // kind per three-body row
    const double *zeta3;
    const double *lam3;
    const double *eta3;
    const double *coef3; // pow(2, 1 - zeta), computed on the host with std::pow
    const int *izeta3;   // zeta as an integer in [0, kMaxIntZeta], or -1
    const int *slot3;
    const int *str3;
    // flat index, in initDescriptor's parameter layout (per block rows x cols,
    // row-major), of each row's first parameter
    const int *pidx2;
    const int *pidx3;
    int n_params;  // sum of rows x cols over the blocks
    int pad2_;
};

// The device table of a parameter set packed on the host, as device_table
// uploads it: every double array, then every int array, at the recorded
// offsets (elements from the start of its block).
struct PackedTable {
    std::vector<double> d;
    std::vector<int> i;
    size_t o_rc = 0, o_a2 = 0, o_b2 = 0, o_z3 = 0, o_l3 = 0, o_e3 = 0, o_c3 = 0;
    size_t o_k2 = 0, o_s2 = 0, o_t2 = 0, o_k3 = 0, o_iz = 0, o_s3 = 0, o_t3 = 0, o_p2 = 0,
           o_p3 = 0;
    // bytes of one upload: the doubles (plus one), then the ints (plus two)
    size_t bytes() const { return sizeof(double) * (d.size() + 1) + sizeof(int) * (i.size() + 2); }
};

}  // namespace ldsf

// Host-side parameter set (opaque in the C-ABI).
struct ldsf_sf {
    std::vector<std::string> species;
    std::vector<double> rcut;  // n_species^2
    std::vector<int> kind, rows, cols, start;
    std::vector<std::vector<double>> par;
    int width = 0;
    int species_resolved = 0;
    int has_three_body = 0;
    int precision = 0;
    // flattened row tables (host copies of SFTab's arrays)
    std::vector<int> k2, slot2, str2, k3, izeta3, slot3, str3, pidx2, pidx3;
    int n_params = 0;
    std::vector<double> a2, b2, zeta3, lam3, eta3, coef3;
    ldsf::PackedTable packed;  // built by finalize_params
    // per-device state, each freed by its own release hook (set by the device
    // runtime, so that this host-side type needs no HIP): the parameter
    // table's upload (kind 0) and the host-pointer calls' staging workspace
    // (kind 1); dev_mu guards the list
    struct Dev {
        int device = -1;
        int kind = 0;
        void *mem = nullptr;
        ldsf::SFTab tab{};
        void (*release)(int device, void *mem) = nullptr;
    };
    std::vector<Dev> dev;
    std::mutex dev_mu;
    ~ldsf_sf();
};

namespace ldsf {
// builds width, start and the flattened row tables from kind/rows/cols/par
int finalize_params(ldsf_sf *sf);
// device table for the current device (uploads on first use)
int device_table(const ldsf_sf *sf, SFTab *out);
// the table's view of an upload of sf->packed at dd (doubles) and di (ints)
SFTab table_view(const ldsf_sf *sf, const double *dd, const int *di);
}  // namespace ldsf

#endif  // LDSF_INTERNAL_H_
