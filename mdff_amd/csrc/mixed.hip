// mixed.hip -- mixed-precision evaluation path.
#include "mixed.h"

namespace mdff {
namespace mixed {

struct MixedState {
  CtxView v;
};

MixedState* create(mdffgpu_ctx*) { return nullptr; }
void destroy(MixedState* m) { delete m; }
void on_cell(MixedState*) {}
void shortrange(MixedState*, int) {}
void es_call(MixedState*, int, const double*, const double*, int, int, int) {}

}  // namespace mixed
}  // namespace mdff
