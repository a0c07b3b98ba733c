// Schur-complement backend (filled in below).
#include "./optimizer.h"

namespace symhip {

std::unique_ptr<SystemBackend> make_schur_backend(const HostProblem&, const SolverOptions&,
                                                  hipStream_t) {
  throw Error("symhip: Schur backend not available yet");
}

}  // namespace symhip
