// The serving loop: eigs / thermal / dynamics solves of spin chains of different size arrive as a
// stream, are admitted against an HBM budget, and run concurrently -- each admitted request on a
// slot with a HIP stream of its own, so the kernels of requests in flight overlap on the device --
// with the sectors kept resident in HBM between requests. Latency is measured from arrival to
// completion. The reference has no serving layer; each request is one of its verbs on one sector
// (src/engine/eigs.cpp, thermal.cpp:161-344, dynamics.cpp:282-690) and the solvers are this
// library's drivers.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace qsa {

enum RequestKind : int { kReqEigs = 0, kReqThermal = 1, kReqDynamics = 2 };

/// One solve on the chain of n_sites spins, periodic: J1 on (i, i+1), J2 on (i, i+2), chi
/// S_i . (S_{i+1} x S_{i+2}) on every (i, i+1, i+2), in the sector n_up = n_sites / 2, momentum 0.
struct Request {
    int kind = kReqEigs;
    int n_sites = 16;
    double j1 = 1.0, j2 = 0.0, chi = 0.0;
    double arrival_s = 0.0;       ///< since the start of the stream
    int k = 4;                    ///< eigs: levels
    int samples = 20, depth = 80; ///< thermal: FTLM samples and Lanczos depth
    std::uint64_t seed = 1;       ///< thermal: the block's seed
    int krylov = 150;             ///< dynamics: continued-fraction depth
};

struct ServeOptions {
    int slots = 8;                           ///< requests in flight
    std::uint64_t hbm_budget = 200ull << 30; ///< admission: reserved bytes of the requests in flight
    bool batched = true;   ///< FTLM samples in one multi-vector apply; false: one sample after another
    bool cache_sectors = true;               ///< keep sectors resident between requests
    bool realtime = true;                    ///< honour arrival times (else every request is there at 0)
    std::uint64_t cache_budget = 48ull << 30;
    std::vector<double> betas;               ///< thermal temperatures (inverse)
    std::vector<double> omega;               ///< dynamics frequency grid
    double eta = 0.05;
    double progress_s = 0.0;   ///< > 0: a line on stderr at most this often as requests complete
};

struct RequestResult {
    std::vector<double> values;  ///< eigs: k levels; thermal: lnZ, E, V per beta; dynamics: S(omega)
    double e0 = 0.0;
    double start_s = 0.0, finish_s = 0.0, latency_s = 0.0;
    double build_s = 0.0;         ///< of start..finish: sectors, programs and CSRs built for it
    std::uint64_t reserved_bytes = 0;
    std::int64_t dim = 0;
    int slot = -1;
    int status = 0;
    std::string error;
};

struct ServeStats {
    double wall_s = 0.0, solves_per_s = 0.0;
    double p50_s = 0.0, p99_s = 0.0, mean_s = 0.0, max_s = 0.0;
    int completed = 0, failed = 0;
    std::uint64_t peak_reserved = 0;
    int peak_in_flight = 0;
    int sector_builds = 0, sector_hits = 0;
    double build_s = 0.0;                      ///< summed over requests
    double service_s[3] = {0.0, 0.0, 0.0};     ///< start..finish summed per kind
};

/// The reserved HBM bytes of a request (sector, CSR, Krylov vectors), before anything is built.
std::uint64_t request_footprint(const Request& r, const ServeOptions& o);

ServeStats serve(const std::vector<Request>& reqs, const ServeOptions& o, std::vector<RequestResult>& out);

}  // namespace qsa
