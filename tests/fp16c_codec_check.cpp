// Runs the device FP16C codec (fluidloom_amd/csrc/fp16c_codec.hpp, host and
// device code) on the CPU for tests/test_fp16c_codec.py.
//
//   fp16c_codec_check decode           -> 65536 float32: decode of every code
//   fp16c_codec_check encode  < x.f32  -> per input: RNE code, then the number
//                                         of the 4096 random values that make
//                                         the stochastic code differ from
//                                         the round-down code, the round-down
//                                         code and the round-up code (5 uint32)
// Binary little-endian on stdin / stdout.
#include <cstdio>
#include <vector>

#include "fp16c_codec.hpp"

using namespace fl::lbm;

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string mode = argv[1];
    if (mode == "decode") {
        std::vector<float> out(65536);
        for (uint32_t c = 0; c < 65536; ++c) out[c] = fp16c_decode((uint16_t)c);
        fwrite(out.data(), sizeof(float), out.size(), stdout);
        return 0;
    }
    if (mode != "encode") return 2;
    float x;
    while (fread(&x, sizeof x, 1, stdin) == 1) {
        const float xs = x * kFp16cDown;
        const uint32_t lo = fp16c_encode_sr_scaled(xs, 0);
        uint32_t hi = lo, up = 0;
        for (uint32_t r = 0; r < 4096; ++r) {
            const uint32_t c = fp16c_encode_sr_scaled(xs, r);
            if (c != lo) {
                if (up && c != hi) return 3;  // more than two results
                hi = c;
                ++up;
            }
        }
        // random bits above the low 12 must not matter
        if (fp16c_encode_sr_scaled(xs, 0xABCDE000u) != lo) return 4;
        const uint32_t rec[5] = {fp16c_encode_rne(x), up, lo, hi, 0};
        fwrite(rec, sizeof(uint32_t), 5, stdout);
    }
    return 0;
}
