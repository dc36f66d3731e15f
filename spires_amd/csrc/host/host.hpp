// Host-side internals of the spires C ABI (no HIP headers here: the device is
// reached only through include/spires_dev.h).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "spires.h"
#include "spires_dev.h"

namespace shost {

constexpr int kMaxMemLen = 20000;   // MAX_MEM_LEN, reference src/neurons/flif_gl.h:4

struct Csr {
    size_t n = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> val;
};

// glibc rand() helpers in the reference's formulas (reservoir.c:308-333).
double urand01();
double gen_weight(double ei);

// init_weights (reservoir.c:345-545): W_in draws then topology, consuming
// rand() in exactly the reference's order; result in CSR (ascending columns,
// zero weights dropped as csr_build_from_dense does).
int generate_topology(size_t n, size_t din, int ctype, double conn, double ei,
                      std::vector<double>& W_in, Csr& W);

// Power iteration (sparse.c:93-127), sequential row dots.
double csr_spectral_radius(const Csr& W);
double dense_spectral_radius(const std::vector<double>& W, size_t n);

void set_error(const std::string& msg);
const std::string& last_error();

}  // namespace shost

struct spires_reservoir {
    spires_reservoir_config cfg{};   // copy; param pointers retained (caller-owned)
    int micro = 1;
    int mode = 0, block_k = 0;
    std::vector<double> W_in;        // [N x Din]
    shost::Csr W;                    // sparse backend (also kept for dense: host copy)
    std::vector<double> dense;       // dense backend [N x N]
    std::vector<double> het_tau;     // PSC-heterogeneous per-edge tau (CSR order or N*N)
    std::vector<double> gl_coef;     // FLIF-GL coefficients made at creation (kept on reset)
    int gl_ncoef = 0;
    spires_dev* dev = nullptr;
};
