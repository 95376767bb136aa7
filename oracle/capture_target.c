/* SPDX-License-Identifier: GPL-3.0-or-later
 * CAPTURE TARGET -- test infrastructure only.  The HEM library that the
 * reference's capture tool (tools/hem_capture.cpp, compiled from its source by
 * `make -C oracle capture`) forwards to while tests/golden/make_capture.py
 * records one batch on a machine without a GPU.  It holds no flash: create
 * returns a token, run reports every cell as not recovered and leaves the
 * states as they are. */
#include <stddef.h>

static int token;

const char* reactive_gpu_hem_numerical_policy_v1(void) { return ""; }

void* reactive_gpu_hem_create_v1(const void* model, size_t bytes, size_t capacity, char* error, size_t size) {
    (void)model; (void)bytes; (void)capacity; (void)error; (void)size;
    return &token;
}

void reactive_gpu_hem_destroy_v1(void* hem) { (void)hem; }

int reactive_gpu_hem_run_v1(void* hem, const double* q, const double* energy, size_t count, void* state, int* success,
                            void* profile, char* error, size_t size) {
    (void)hem; (void)q; (void)energy; (void)state; (void)profile; (void)error; (void)size;
    for (size_t c = 0; c < count; ++c) success[c] = 0;
    return 0;
}

int reactive_gpu_hem_run_v2(void* hem, const double* q, const double* energy, const void* capillary, size_t count,
                            void* state, int* success, void* profile, char* error, size_t size) {
    (void)capillary;
    return reactive_gpu_hem_run_v1(hem, q, energy, count, state, success, profile, error, size);
}

int reactive_gpu_hem_initialize_tp_v1(void* hem, double* q, double* energy, const void* capillary, size_t count,
                                      void* state, int* success, void* profile, char* error, size_t size) {
    (void)hem; (void)q; (void)energy; (void)capillary; (void)count; (void)state; (void)success; (void)profile;
    (void)error; (void)size;
    return 1;
}
