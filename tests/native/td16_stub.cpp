// Stub of the few C-ABI entry points the JNI shim calls, for the CPU build of
// tests/native/jni_driver.cpp: it records what the shim passed and answers
// by the number of symbols (3: the witness (0, 1, 2); 4: no witness; 5: an
// error with a message), so the shim's marshalling is checked without a GPU.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/td16_abi.h"

int gStubSlots = 0, gStubDevices[8] = {0}, gStubNdev = 0;
long long gStubBudget = 0;
uint8_t gStubBeta[32] = {0};
int gStubN = 0, gStubServeQueue = 0, gStubServeSlots = 0;

namespace {
int fail(char *msg, int cap, const char *text) {
  if (msg && cap > 0) std::snprintf(msg, (size_t)cap, "%s", text);
  return TD16_INVALID_ARGUMENT;
}
}  // namespace

extern "C" {

int td16_init_slots(int total, const int32_t *devices, int ndev, int64_t budget, char *msg,
                    int cap) {
  gStubSlots = total;
  gStubNdev = ndev;
  for (int q = 0; q < ndev && q < 8; ++q) gStubDevices[q] = devices[q];
  gStubBudget = budget;
  (void)msg, (void)cap;
  return TD16_OK;
}

int td16_search(int slot, const uint8_t *beta, const uint8_t *omega, int n, int initial_even,
                float min_rate, int max_moves, int full_sorting, int32_t *out_positions,
                int32_t *out_len, char *msg, int cap) {
  (void)slot, (void)omega, (void)initial_even, (void)min_rate, (void)max_moves,
      (void)full_sorting;
  gStubN = n;
  std::memcpy(gStubBeta, beta, (size_t)n);
  if (n == 5) return fail(msg, cap, "stub rejects five symbols");
  *out_len = 0;
  if (n == 3) {
    out_positions[0] = 0, out_positions[1] = 1, out_positions[2] = 2;
    *out_len = 1;
  }
  return TD16_OK;
}

int td16_serve_start(const td16_serve_options *opt, char *msg, int cap) {
  (void)msg, (void)cap;
  gStubServeSlots = opt->slots;
  gStubServeQueue = opt->queue_capacity;
  return TD16_OK;
}

int td16_serve_search(const td16_request *req, td16_result *res, char *msg, int cap) {
  std::memset(res, 0, sizeof(*res));
  if (req->n == 5) return fail(msg, cap, "stub rejects five symbols");
  if (req->n == 3) {
    res->found = 1;
    res->depth = 1;
    res->moves[0][0] = 0, res->moves[0][1] = 1, res->moves[0][2] = 2;
  }
  return TD16_OK;
}

}  // extern "C"
