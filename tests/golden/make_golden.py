"""Write tests/golden/ fixtures from the reference's own CPU path.

Run in a container that has the reference sources (after `make -C oracle ref`):

    python tests/golden/make_golden.py

The compiled reference (oracle/_ref/libnbody_ref.so, built from
src/nbody_CPU_*.cpp and src/nbody_util.cpp where they lie) produces:

  ref_seed7_n1024.npz   seedRandom(7) + randomUnitBodies(N=1024, 0.9, 8.0)
                        (src/nbody.cu:919,996) and the forces of CPU_SOA,
                        CPU_AOS and AVX at softeningSquared = 0.1f*0.1f;
  ref_seed7_n4096.npz   the same ICs at N=4096 and the tiled forces
                        (CPU_SOA_tiled / CPU_AOS_tiled need N % 4096 == 0);
  ref_digests.json      SHA-256 of the IC and force bytes at N = 16384 and
                        98304 (seed 7), too large to commit as arrays.

Only inputs and outputs are written; nothing of the reference's text.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import ref_cpu  # noqa: E402

EPS2 = np.float32(0.1) * np.float32(0.1)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    r = ref_cpu()
    if r is None:
        sys.exit("oracle/_ref/libnbody_ref.so missing: run `make -C oracle ref` first")
    pm, vm = r.random_unit_bodies(1024, 7)
    out = {"pos_mass": pm, "vel_inv_mass": vm}
    out["force_cpu_soa"] = r.cpu_soa(pm, EPS2)[0][:, :3]
    out["force_cpu_aos"] = r.cpu_aos(pm, EPS2)[0][:, :3]
    out["force_cpu_avx"] = r.cpu_avx(pm, EPS2)[0][:, :3]
    np.savez_compressed(os.path.join(HERE, "ref_seed7_n1024.npz"), **out)

    pm, vm = r.random_unit_bodies(4096, 7)
    out = {"pos_mass": pm, "vel_inv_mass": vm}
    out["force_cpu_soa"] = r.cpu_soa(pm, EPS2)[0][:, :3]
    out["force_cpu_aos"] = r.cpu_aos(pm, EPS2)[0][:, :3]
    out["force_cpu_soa_tiled"] = r.cpu_soa_tiled(pm, EPS2)[0][:, :3]
    out["force_cpu_aos_tiled"] = r.cpu_aos_tiled(pm, EPS2)[0][:, :3]
    np.savez_compressed(os.path.join(HERE, "ref_seed7_n4096.npz"), **out)

    digests = {}
    for n in (16384, 98304):
        pm, vm = r.random_unit_bodies(n, 7)
        d = {"pos_mass": sha(pm), "vel_inv_mass": sha(vm)}
        if n == 16384:
            d["force_cpu_soa"] = sha(r.cpu_soa(pm, EPS2)[0][:, :3])
            d["force_cpu_aos"] = sha(r.cpu_aos(pm, EPS2)[0][:, :3])
        digests[str(n)] = d
    with open(os.path.join(HERE, "ref_digests.json"), "w") as f:
        json.dump({"seed": 7, "scale": 0.9, "velscale": 8.0, "softening_squared": float(EPS2),
                   "host_build": "g++ -O3 -ffast-math -march=x86-64-v4 -fopenmp",
                   "digests": digests}, f, indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
