// Host build of the device field templates (sumcheck-triton_amd/csrc/field.hpp), so the
// arithmetic can be checked against Python integers without a GPU. Test-only.
#include <stdint.h>

#include "field.hpp"

using namespace sc;

template <class F>
static void run(int op, uint64_t n, const uint32_t* a, const uint32_t* b, uint32_t* out) {
  constexpr int L = F::L;
  for (uint64_t i = 0; i < n; ++i) {
    Elem<L> x, y, z;
    for (int k = 0; k < L; ++k) {
      x.v[k] = a[i * L + k];
      y.v[k] = b[i * L + k];
    }
    if (op == 0) z = mont_mul<F>(x, y);
    else if (op == 1) z = add<F>(x, y);
    else if (op == 2) z = sub<F>(x, y);
    else if (op == 3) z = mont_mul_cios<F>(x, y);
    else z = mont_mul_fips<F>(x, y);
    for (int k = 0; k < L; ++k) out[i * L + k] = z.v[k];
  }
}

extern "C" void field_host_op(int bit_width, int op, uint64_t n, const uint32_t* a, const uint32_t* b,
                              uint32_t* out) {
  switch (bit_width) {
    case 32: run<FieldP32>(op, n, a, b, out); break;
    case 64: run<FieldP64>(op, n, a, b, out); break;
    case 128: run<FieldP128>(op, n, a, b, out); break;
    default: run<FieldP256>(op, n, a, b, out); break;
  }
}
