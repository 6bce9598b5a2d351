"""Write the golden fixtures of tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_push (the reference's particle.h, Particle::Move) and
oracle/_ref/ref_rng (the reference's maxwell.cxx: g05cae_/g05dde_, and
writeParamsFile for params_ref.dat), both
built by `make -C oracle ref` from the sources in the reference checkout,
and stores their inputs and outputs as .npz data.  Run where that checkout
exists:  python tests/golden/gen_golden.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def boris_vectors(n=1024, seed=7):
    rng = np.random.default_rng(seed)
    E = rng.normal(0, 0.3, (n, 3))
    H = rng.normal(0, 0.8, (n, 3))
    x = rng.uniform(0.0, 3.2, (n, 3))
    p = rng.normal(0, 0.4, (n, 3))
    q_m = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0 / 1836.0)
    tau = np.full(n, 0.05)
    # special cases: zero fields, zero momentum, strong field, relativistic
    E[:8] = 0.0
    H[8:16] = 0.0
    p[16:24] = 0.0
    p[24:32] *= 30.0
    H[32:40] *= 50.0
    rec = np.concatenate([E, H, x, p, q_m[:, None], tau[:, None]], axis=1)
    inp = np.int64(n).tobytes() + rec.astype("<f8").tobytes()
    out = subprocess.run([os.path.join(REF_BIN, "ref_push")], input=inp, stdout=subprocess.PIPE,
                         check=True).stdout
    res = np.frombuffer(out, dtype="<f8").reshape(n, 6)
    return dict(E=E, H=H, x=x, p=p, q_m=q_m, tau=tau, p_out=res[:, :3], x1_out=res[:, 3:])


def rng_vectors(nu=2000, nn=2000, mean=0.3, sd=0.005):
    out = subprocess.run([os.path.join(REF_BIN, "ref_rng"), str(nu), str(nn), repr(mean),
                          repr(sd)], stdout=subprocess.PIPE, check=True).stdout
    v = np.frombuffer(out, dtype="<f8")
    return dict(uniform=v[:nu], normal=v[nu:], mean=np.float64(mean), sd=np.float64(sd))


def params_file():
    """000_params.dat written by the reference's writeParamsFile."""
    subprocess.run([os.path.join(REF_BIN, "ref_rng"), "params"], cwd=HERE, check=True)
    os.replace(os.path.join(HERE, "000_params.dat"), os.path.join(HERE, "params_ref.dat"))


def main():
    if not os.path.exists(os.path.join(REF_BIN, "ref_push")):
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference exists")
    np.savez_compressed(os.path.join(HERE, "boris_ref.npz"), **boris_vectors())
    np.savez_compressed(os.path.join(HERE, "rng_ref.npz"), **rng_vectors())
    params_file()
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
