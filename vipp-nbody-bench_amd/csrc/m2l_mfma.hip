// m2l_mfma.hip - M2L on the matrix cores (split BF16 operands); see fmm_m2l_split.
#include "gravity.hpp"

namespace vg {

void fmm_m2l_split(Ctx &) { throw Error(1, "split-precision M2L is not built in this library"); }

}  // namespace vg
