"""Writes tests/golden/philox_vectors.json.

Two independent sources pin the Philox4x32-10 used by the oracle and the
kernels:
  - the Random123 known-answer vectors for philox4x32_10 (kat_vectors of
    Random123 1.09, published with Salmon et al. SC'11);
  - rocRAND's own implementation (rocrand_philox4x32_10.h, header source in
    /opt/rocm/include), compiled here for the host and evaluated on a grid
    of counters and keys.
Run from the repository root: python tests/golden/gen_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

KAT = [
    # counter (4), key (2), expected output (4)
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]

SRC = r"""
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main() {
    Engine eng;
    uint32_t s = 12345u;
    auto nxt = [&]() { s = s * 1664525u + 1013904223u; return s; };
    printf("[");
    for (int i = 0; i < 256; ++i) {
        uint4 c;
        uint2 k;
        if (i < 64) { c = uint4{(uint32_t)i, 7u, 3u, 0u}; k = uint2{(uint32_t)(i * 977u), 20251203u}; }
        else { c = uint4{nxt(), nxt(), nxt(), nxt()}; k = uint2{nxt(), nxt()}; }
        uint4 o = eng.ten_rounds(c, k);
        printf("%s[[%u,%u,%u,%u],[%u,%u],[%u,%u,%u,%u]]", i ? "," : "", c.x, c.y, c.z, c.w, k.x, k.y,
               o.x, o.y, o.z, o.w);
    }
    printf("]\n");
    return 0;
}
"""


def rocrand_vectors():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["hipcc", "-O1", "-x", "hip", "--offload-arch=gfx950", src, "-o", exe],
                       check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def main():
    data = {"random123_kat": KAT, "rocrand": rocrand_vectors()}
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump(data, f)
    print("wrote", len(data["rocrand"]), "rocRAND vectors")


if __name__ == "__main__":
    main()
