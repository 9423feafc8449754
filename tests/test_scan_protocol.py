"""The single-pass scan's aggregate-word protocol (diso_amd/csrc/diso_scan_word.h, the header the
count kernel uses), compiled with g++ and driven on the CPU.

A model control block (two word buffers, an epoch) runs a sequence of calls of a few blocks each.
In one call a block is late: the call's last block gives up on it (the scan timeout) and advances
the epoch; the late block then writes its word, after the next call has already cleared that
buffer. The call after that reuses the buffer, and its readers must not take the stale word for
the current call's: it reads as not published, so the reader counts that range itself, and every
call's prefixes stay exact. Also: the word's fields round-trip at their limits, an empty word is
never published, and tags differ for calls that share a buffer.
"""
import os
import subprocess


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "diso_scan_word.h"
using namespace diso_scan;

#define CHECK(c) do { if (!(c)) { printf("FAIL line %d: %s\n", __LINE__, #c); return 1; } } while (0)

enum { NB = 6, CALLS = 12 };
static uint64_t buf[2][NB];
static uint32_t dirty[2];
static uint32_t epoch;

static int64_t truth_v(int call, int b) { return 1000 * call + 10 * b + 1; }
static int64_t truth_f(int call, int b) { return 7 * call + b; }

int main() {
    // fields at their limits
    const int64_t vmax = (1ll << kCountBits) - 1, fmax = (1ll << kFaceBits) - 1;
    for (uint32_t e = 0; e < 7; ++e) {
        uint64_t w = pack(vmax, fmax, 7, tag_of(e));
        CHECK(published(w, tag_of(e)) && count_v(w) == vmax && count_f(w) == fmax && flags_of(w) == 7);
        CHECK(!published(0, tag_of(e)));
        CHECK(tag_of(e) != tag_of(e + 2) && tag_of(e) != tag_of(e + 4));  // calls sharing a buffer
        CHECK(buffer_of(e) == buffer_of(e + 2) && buffer_of(e) != buffer_of(e + 1));
    }
    int late_block = 3, late_call = 4;
    uint64_t late_word = 0;
    for (int call = 0; call < CALLS; ++call) {
        const uint32_t ep = epoch, p = buffer_of(ep);
#ifdef CONSTANT_TAG
        const uint32_t tag = 1;  // a protocol without epochs, for the test's own sensitivity check
#else
        const uint32_t tag = tag_of(ep);
#endif
        // every block clears its share of the other buffer (the extent the earlier call wrote)
        for (uint32_t i = 0; i < dirty[p ^ 1u]; ++i) buf[p ^ 1u][i] = 0;
        // the late block of the timed-out call writes its word only now, after that clear
        if (call == late_call + 1) buf[buffer_of(ep - 1)][late_block] = late_word;
        // blocks publish; the late block of late_call does not in time, nor (merely slow) the same
        // block two calls later, whose readers then find the stale word in its slot
        const bool slow = call == late_call + 2;
        for (int b = 0; b < NB; ++b)
            if (!((call == late_call || slow) && b == late_block))
                buf[p][b] = pack(truth_v(call, b), truth_f(call, b), 1, tag);
        // every block's exclusive prefix from its predecessors' words; unpublished ranges recounted
        for (int b = 0; b < NB; ++b) {
            int64_t pv = 0, pf = 0, wv = 0, wf = 0;
            for (int j = 0; j < b; ++j) {
                const uint64_t w = buf[p][j];
                if (published(w, tag)) { pv += count_v(w); pf += count_f(w); }
                else { pv += truth_v(call, j); pf += truth_f(call, j); }  // the reader counts range j
                wv += truth_v(call, j); wf += truth_f(call, j);
            }
            CHECK(pv == wv && pf == wf);
        }
        if (slow) buf[p][late_block] = pack(truth_v(call, late_block), truth_f(call, late_block), 1, tag);
        if (call == late_call) late_word = pack(truth_v(call, late_block), truth_f(call, late_block), 1, tag);
        // the last block records the extent and advances the epoch (on time or after giving up)
        dirty[p] = NB;
        dirty[p ^ 1u] = 0;
        epoch = ep + 1;
    }
    printf("OK\n");
    return 0;
}
"""


def _build_and_run(tmp_path, defines):
    src = tmp_path / "scan_word_test.cpp"
    exe = tmp_path / ("scan_word_test" + "".join(defines))
    src.write_text(PROGRAM)
    cc = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", *[f"-D{d}" for d in defines], "-I",
                         os.path.join(ROOT, "diso_amd", "csrc"), str(src), "-o", str(exe)], capture_output=True,
                        text=True)
    assert cc.returncode == 0, cc.stderr
    return subprocess.run([str(exe)], capture_output=True, text=True)


def test_stale_words_of_a_timed_out_call_never_count(tmp_path):
    r = _build_and_run(tmp_path, [])
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr


def test_the_model_detects_a_protocol_without_epochs(tmp_path):
    """The same scenario with one constant tag: the stale word is taken for the current call's and
    a prefix comes out wrong, so the model above does exercise the stale-word case."""
    r = _build_and_run(tmp_path, ["CONSTANT_TAG"])
    assert r.returncode == 1 and "FAIL" in r.stdout
