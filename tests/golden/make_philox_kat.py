"""Write tests/golden/philox_kat.json from rocRAND's Philox4x32-10 engine.

rocRAND (a ROCm library shipped as a header in this image) is an independent
implementation of Philox4x32-10.  This script compiles a small host program
against its header, evaluates the engine's ten_rounds() on chosen counters and
keys, and stores the results; test_oracle_rng.py checks the oracle against them.

    python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  probe p;
  unsigned cases[][6] = {
    {0,0,0,0,0,0}, {1,0,0,0,0,0}, {0,1,0,0,0,0}, {0,0,1,0,0,0}, {0,0,0,1,0,0},
    {0,0,0,0,1,0}, {0,0,0,0,0,1}, {123456789u,42u,3u,0u,0x10u,0u},
    {7u,1000u,2u,0u,0u,0x40u}, {0xdeadbeefu,0xcafebabeu,0x01234567u,0x89abcdefu,0xfeedfaceu,0x0badf00du},
    {4294967295u,4294967295u,4294967295u,4294967295u,4294967295u,4294967295u},
  };
  for (auto& c : cases) {
    uint4 r = p.run(uint4{c[0],c[1],c[2],c[3]}, uint2{c[4],c[5]});
    printf("%u %u %u %u %u %u %u %u %u %u\n", c[0],c[1],c[2],c[3],c[4],c[5], r.x,r.y,r.z,r.w);
  }
  return 0;
}
"""


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "philox_kat.json")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "kat.cpp")
        exe = os.path.join(d, "kat")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["hipcc", "-O1", "-o", exe, src], check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    vec = []
    for ln in lines:
        if not ln.strip():
            continue
        v = [int(t) for t in ln.split()]
        vec.append({"ctr": v[0:4], "key": v[4:6], "out": v[6:10]})
    with open(out, "w") as f:
        json.dump({"source": "rocRAND philox4x32_10_engine::ten_rounds (ROCm 7.2 header)",
                   "vectors": vec}, f, indent=1)
    print("wrote", out, len(vec), "vectors")


if __name__ == "__main__":
    main()
