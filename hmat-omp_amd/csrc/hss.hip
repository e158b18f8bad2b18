// HSS engine on gfx950 (placeholder until the nested-basis build lands).
#include "hss.hpp"

namespace hmat {

HssEngine::HssEngine(const hmat_options& o) : opt(o) {
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  own_stream = true;
}
HssEngine::~HssEngine() {
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}
void HssEngine::build(const GenDesc&, int64_t, int, bool, const double*, int64_t) {
  throw Error(HMAT_ERR_STATE, "HSS build not available yet");
}
void HssEngine::matvec(const double*, int64_t, int, double*, int64_t) {
  throw Error(HMAT_ERR_STATE, "HSS not built");
}
void HssEngine::factor() { throw Error(HMAT_ERR_STATE, "HSS not built"); }
void HssEngine::solve(const double*, int64_t, int, double*, int64_t) {
  throw Error(HMAT_ERR_STATE, "HSS not built");
}
void HssEngine::to_dense_host(double*, int64_t) { throw Error(HMAT_ERR_STATE, "HSS not built"); }
int HssEngine::root_rank() const { return 0; }
int64_t HssEngine::omega_length(const hmat_options&, int64_t, int) { return 0; }

}  // namespace hmat
