// runtime_validation.cpp -- TEST INFRASTRUCTURE: drives the host side of the C
// ABI (menodax_amd/csrc/runtime.cpp, include/menodax_hip.h) through every path
// that runs without a GPU -- kernel registration and the launchers' argument
// validation, which rejects malformed launches before any HIP call -- so that a
// build under AddressSanitizer and UndefinedBehaviorSanitizer (Makefile beside
// this file) checks them. Exit status 0 when every call returns what the header
// promises; the sanitizers abort on their own findings.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "menodax_hip.h"

static int failures = 0;

static void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "runtime validation: %s failed (last error: %s)\n", what,
                 mdx_last_error());
    ++failures;
  }
}

static bool error_says(int32_t rc, const char* fragment) {
  return rc != 0 && std::strstr(mdx_last_error(), fragment) != nullptr;
}

int main() {
  expect(mdx_abi_version() == 1, "abi version");
  expect(std::strlen(mdx_last_error()) == 0, "no error before the first call");

  // kernels are registered from any bytes; the code object is loaded on first launch
  std::vector<char> code(4096, 0x5a);
  int64_t plain = 0, refill = 0, adjoint = 0;
  expect(mdx_kernel_create(code.data(), code.size(), "mdx_rodas5p", 64, 0, &plain) == 0,
         "register a plain kernel");
  expect(mdx_kernel_create(code.data(), code.size(), "mdx_rodas5p_refill", 64,
                           MDX_KERNEL_PERSISTENT, &refill) == 0,
         "register a persistent kernel");
  expect(mdx_kernel_create(code.data(), code.size(), "mdx_rodas5p_adjoint", 64,
                           MDX_KERNEL_ADJOINT | MDX_KERNEL_UNIT_WEIGHTS, &adjoint) == 0,
         "register an adjoint kernel");
  int64_t h = 0;
  expect(error_says(mdx_kernel_create(nullptr, 16, "k", 64, 0, &h), "null argument"),
         "null code");
  expect(error_says(mdx_kernel_create(code.data(), code.size(), "k", 0, 0, &h), "block"),
         "block 0");
  expect(error_says(mdx_kernel_create(code.data(), code.size(), "k", 64,
                                      MDX_KERNEL_ADJOINT | MDX_KERNEL_PERSISTENT, &h),
                    "not a persistent kernel"),
         "adjoint and persistent");
  expect(mdx_workspace_bytes(refill, 1000) == 256 && mdx_workspace_bytes(plain, 1000) == 0,
         "workspace sizes");
  expect(mdx_workspace_bytes(123456, 10) == -1, "workspace of an unknown handle");
  expect(mdx_adjoint_workspace_bytes(-1, 3, 8, 2) == -1, "adjoint workspace, bad extents");

  // launches rejected before the GPU
  double buf[64] = {0};
  int32_t ibuf[16] = {0};
  auto launch = [&](int64_t k, int64_t n, int32_t nv, int32_t ns, int32_t layout,
                    const double* y0, const double* w, void* ws, int64_t wsb, int32_t grid) {
    return mdx_solve_launch(k, nullptr, grid, n, nv, 3, ns, layout, y0, buf, buf, 0.0, 1e-6,
                            1e-8, 100, w, buf, ibuf, ibuf, ibuf, nullptr, 0, ws, wsb);
  };
  expect(error_says(launch(999, 4, 3, 5, 1, buf, buf, nullptr, 0, 0), "unknown kernel handle"),
         "unknown handle");
  expect(error_says(launch(plain, -1, 3, 5, 1, buf, buf, nullptr, 0, 0), "bad extents"),
         "negative N");
  expect(error_says(launch(plain, 4, 0, 5, 1, buf, buf, nullptr, 0, 0), "bad extents"),
         "no variables");
  expect(error_says(launch(plain, 4, 3, 5, 9, buf, buf, nullptr, 0, 0), "layout"), "layout");
  expect(error_says(launch(plain, 4, 3, 5, 1, nullptr, buf, nullptr, 0, 0), "null buffer"),
         "null y0");
  expect(error_says(launch(plain, 4, 3, 5, 1, buf, nullptr, nullptr, 0, 0), "weights"),
         "null weights");
  expect(error_says(launch(plain, 200, 3, 5, 1, buf, buf, nullptr, 0, 3), "grid"), "small grid");
  expect(error_says(launch(refill, 4, 3, 5, 1, buf, buf, nullptr, 0, 0), "workspace"),
         "persistent without workspace");
  expect(error_says(launch(adjoint, 4, 3, 5, 1, buf, buf, nullptr, 0, 0), "adjoint"),
         "adjoint through the solve launcher");
  int64_t order[4] = {0, 1, 2, 3};
  expect(error_says(mdx_solve_launch_ordered(plain, nullptr, 0, 4, 3, 3, 5, 1, buf, buf, buf,
                                             0.0, 1e-6, 1e-8, 100, buf, buf, ibuf, ibuf, ibuf,
                                             nullptr, 0, nullptr, 0, order),
                    "persistent"),
         "claim order on a plain kernel");
  expect(error_says(mdx_cost_pilot_launch(plain, nullptr, 4, 3, 3, 5, 1, buf, buf, buf, 0.0,
                                          1e-6, 1e-8, 0, buf, buf, ibuf, ibuf, ibuf, buf),
                    "steps"),
         "pilot without steps");
  expect(launch(plain, 0, 3, 5, 1, buf, buf, nullptr, 0, 0) == 0, "an empty ensemble is a no-op");

  // the adjoint launchers
  auto adj = [&](int64_t k, int64_t n, void* ws, int64_t wsb) {
    return mdx_adjoint_launch(k, nullptr, n, 3, 3, 5, 1, buf, buf, buf, 0.0, 1e-6, 1e-8, 100,
                              nullptr, buf, buf, buf, ibuf, 8, 2, ws, wsb);
  };
  expect(error_says(adj(plain, 4, buf, 8), "not an adjoint kernel"), "adjoint of a plain kernel");
  expect(error_says(adj(adjoint, 4, buf, 8), "workspace holds"), "small workspace");
  expect(error_says(mdx_adjoint_launch_static(adjoint, nullptr, int64_t(1) << 30, 3, 3, 5, 1,
                                              buf, buf, buf, 0.0, 1e-6, 1e-8, 100, nullptr, buf,
                                              buf, buf, ibuf, 8, 2),
                    "static scratch"),
         "static adjoint beyond its scratch");
  expect(adj(adjoint, 0, buf, 0) == 0, "an empty adjoint is a no-op");

  std::printf("runtime validation: %s\n", failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}
