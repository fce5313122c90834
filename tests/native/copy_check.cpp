// Native check of the frame copy pool (pyvr_amd/csrc/hvr_copy.h), built with
// ThreadSanitizer by tests/test_sanitizers.py: copies of sizes from 0 to
// 9 MiB (below and above the 1 MiB split, not multiples of the part size)
// at unaligned offsets equal memcpy's, from one caller and from four
// callers at once, and in a child forked after the workers started (where
// copies stay on the calling thread).  Exit status 0 on success.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/wait.h>
#include <unistd.h>

#include <thread>
#include <vector>

#include "hvr_copy.h"

static uint64_t g_state = 0x9e3779b97f4a7c15ull;
static uint64_t next()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}

static bool one(size_t n, size_t so, size_t doff, uint8_t salt)
{
    std::vector<uint8_t> src(n + so + 1), dst(n + doff + 1, 0xAB), ref(n + doff + 1, 0xAB);
    for (size_t i = 0; i < src.size(); ++i) src[i] = (uint8_t)(i * 131u + salt);
    memcpy(ref.data() + doff, src.data() + so, n);
    hvr::CopyPool::get().copy(dst.data() + doff, src.data() + so, n);
    return dst == ref;
}

int main()
{
    int fails = 0;
    const size_t sizes[] = {0, 1, 4095, (size_t)1 << 20, ((size_t)1 << 20) + 3, (size_t)4 << 20, 4 * 1024 * 1024 + 17,
                            9 * 1024 * 1024 - 5};
    for (size_t n : sizes)
        for (int k = 0; k < 3; ++k)
            if (!one(n, next() % 64, next() % 64, (uint8_t)k)) {
                fprintf(stderr, "FAIL size %zu\n", n);
                ++fails;
            }
    std::vector<std::thread> callers;
    std::vector<int> bad(4, 0);
    for (int t = 0; t < 4; ++t)
        callers.emplace_back([t, &bad] {
            for (int k = 0; k < 6; ++k)
                if (!one(((size_t)1 << 20) + (size_t)(t * 7919 + k * 104729), (size_t)(t + k) % 64, (size_t)k % 64,
                         (uint8_t)(t * 16 + k)))
                    ++bad[t];
        });
    for (auto &c : callers) c.join();
    for (int t = 0; t < 4; ++t) fails += bad[t];
    const pid_t child = fork();
    if (child == 0) _exit(one((size_t)3 << 20, 5, 7, 99) && one((size_t)1 << 20, 0, 0, 98) ? 0 : 1);
    int status = 0;
    if (child < 0 || waitpid(child, &status, 0) != child || !WIFEXITED(status) || WEXITSTATUS(status) != 0) {
        fprintf(stderr, "FAIL forked child\n");
        ++fails;
    }
    if (fails) return 1;
    printf("copy ok\n");
    return 0;
}
