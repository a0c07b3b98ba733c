// Compile-time description of each registered factor type for the device kernels.
#pragma once

#include "../../include/symhip_types.h"
#include "./common.h"
#include "./geo.h"

namespace symhip {

template <int Code>
struct FT;

#define SYMHIP_FT(CODE, D_, T_, NA_, NL_, TD0, TD1, TD2, BODY)                        \
  template <>                                                                        \
  struct FT<CODE> {                                                                  \
    static constexpr int D = D_, T = T_, NA = NA_, NL = NL_;                         \
    static constexpr int td(int i) { return i == 0 ? TD0 : (i == 1 ? TD1 : TD2); }    \
    static constexpr int foff(int i) { return i == 0 ? 0 : (i == 1 ? TD0 : TD0 + TD1); } \
    template <typename S>                                                            \
    static SYMHIP_HD void eval(const double* const* a, S* r, S (*J)[T_]) { BODY; }   \
  };

SYMHIP_FT(SYMHIP_FACTOR_PRIOR_POSE3, 6, 6, 4, 1, 6, 0, 0, factor_prior_pose3<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_BETWEEN_POSE3, 6, 12, 5, 2, 6, 6, 0, factor_between_pose3<S>(a, r, J, 0))
SYMHIP_FT(SYMHIP_FACTOR_PRIOR_ROT3, 3, 3, 4, 1, 3, 0, 0, factor_prior_rot3<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_BETWEEN_ROT3, 3, 6, 5, 2, 3, 3, 0, factor_between_rot3<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_SNAVELY, 2, 12, 5, 3, 6, 3, 3, factor_snavely<S>(a, r, J, 0))
SYMHIP_FT(SYMHIP_FACTOR_SNAVELY_HUBER, 2, 12, 6, 3, 6, 3, 3, factor_snavely<S>(a, r, J, 1))
SYMHIP_FT(SYMHIP_FACTOR_REPROJ_CAUCHY, 2, 9, 6, 2, 6, 3, 0, factor_reproj_cauchy<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_MATCHING, 3, 6, 4, 1, 6, 0, 0, factor_matching<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_ODOMETRY, 6, 12, 5, 2, 6, 6, 0, factor_odometry<S>(a, r, J))
SYMHIP_FT(SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY, 6, 12, 6, 2, 6, 6, 0,
          factor_between_pose3<S>(a, r, J, 1))
SYMHIP_FT(SYMHIP_FACTOR_BARRON_VECTOR5, 5, 5, 4, 1, 5, 0, 0, factor_barron5<S>(a, r, J))

#undef SYMHIP_FT

// Calls F.template operator()<FT>() for the runtime factor type code.
template <typename F>
inline void dispatch_factor_type(int code, F&& f) {
  switch (code) {
    case SYMHIP_FACTOR_PRIOR_POSE3: f.template operator()<SYMHIP_FACTOR_PRIOR_POSE3>(); break;
    case SYMHIP_FACTOR_BETWEEN_POSE3: f.template operator()<SYMHIP_FACTOR_BETWEEN_POSE3>(); break;
    case SYMHIP_FACTOR_PRIOR_ROT3: f.template operator()<SYMHIP_FACTOR_PRIOR_ROT3>(); break;
    case SYMHIP_FACTOR_BETWEEN_ROT3: f.template operator()<SYMHIP_FACTOR_BETWEEN_ROT3>(); break;
    case SYMHIP_FACTOR_SNAVELY: f.template operator()<SYMHIP_FACTOR_SNAVELY>(); break;
    case SYMHIP_FACTOR_SNAVELY_HUBER: f.template operator()<SYMHIP_FACTOR_SNAVELY_HUBER>(); break;
    case SYMHIP_FACTOR_REPROJ_CAUCHY: f.template operator()<SYMHIP_FACTOR_REPROJ_CAUCHY>(); break;
    case SYMHIP_FACTOR_MATCHING: f.template operator()<SYMHIP_FACTOR_MATCHING>(); break;
    case SYMHIP_FACTOR_ODOMETRY: f.template operator()<SYMHIP_FACTOR_ODOMETRY>(); break;
    case SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY:
      f.template operator()<SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY>();
      break;
    case SYMHIP_FACTOR_BARRON_VECTOR5: f.template operator()<SYMHIP_FACTOR_BARRON_VECTOR5>(); break;
    default: throw Error("symhip: unknown factor type");
  }
}

}  // namespace symhip
