"""Writes motion_generators.npz: input/output vectors of the reference's
seeded NumPy generators, for the bit-exact tests in tests/test_generators.py.

The reference modules import CuPy, SigPy, matplotlib and nibabel at the top,
none of which is installed here, so the four functions are taken from the
module sources by name (ast) and run with NumPy alone:
  simulation.py:99-102   generate_complex_gaussian_noise (xp = numpy)
  simulation.py:104-174  generate_motion_parameters_new
  simulation.py:24-63    sample_disorder
  utils.py:5-36          generate_motion_parameters
and writes disorder_orders.npz: orders of the reference's DISORDER sampler
(simulation.py:192-245, DisorderConfig and make_disorder_order, run with the
standard library's random, math.ceil and dataclasses) for the tile and
acceleration settings the reconstruction workloads use.
Only the outputs are stored.  Run from the repository root:
  python tests/golden/make_generator_vectors.py [reference dir]
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_functions(path, names, extra=None):
    tree = ast.parse(open(path).read(), filename=path)
    keep = [n for n in tree.body if isinstance(n, (ast.FunctionDef, ast.ClassDef)) and n.name in names]
    missing = set(names) - {n.name for n in keep}
    if missing:
        raise SystemExit(f"{path}: functions not found: {sorted(missing)}")
    ns = {"np": np, **(extra or {})}
    exec(compile(ast.Module(body=keep, type_ignores=[]), path, "exec"), ns)
    return ns


def main(ref):
    sim = load_functions(os.path.join(ref, "simulation.py"),
                         ["generate_complex_gaussian_noise", "generate_motion_parameters_new", "sample_disorder"])
    utl = load_functions(os.path.join(ref, "utils.py"), ["generate_motion_parameters"])
    out = {}
    # generate_motion_parameters_new: (num_states, low_freq_var, high_freq_var, spike_prob, seed, rot_limit_deg)
    new_cases = [(64, 1, 20.0, 0.02, 123456789, 20.0), (128, 0.1, 5.0, 0.02, 123456789, 20.0),
                 (16, 4, 20.0, 0.2, 7, 20.0), (200, 50, 80.0, 0.01, 1, 3.0), (1, 1, 20.0, 0.02, 5, 20.0)]
    for i, (n, lf, hf, sp, sd, lim) in enumerate(new_cases):
        out[f"new_{i}_args"] = np.array([n, lf, hf, sp, sd, lim], dtype=np.float64)
        out[f"new_{i}"] = sim["generate_motion_parameters_new"](n, low_freq_var=lf, high_freq_var=hf, spike_prob=sp,
                                                                 seed=sd, rot_limit_deg=lim)
    # generate_motion_parameters: (num_states, low_freq_var, high_freq_var, spike_prob, seed)
    old_cases = [(64, 0.1, 5.0, 0.02, 123456789), (32, 1.0, 20.0, 0.3, 11), (1, 0.5, 1.0, 0.5, 3)]
    for i, (n, lf, hf, sp, sd) in enumerate(old_cases):
        out[f"old_{i}_args"] = np.array([n, lf, hf, sp, sd], dtype=np.float64)
        out[f"old_{i}"] = utl["generate_motion_parameters"](n, low_freq_var=lf, high_freq_var=hf, spike_prob=sp,
                                                             seed=sd)
    # generate_complex_gaussian_noise: real std, complex std (as simulation.py:445-447 passes max(ksp))
    out["noise_0"] = sim["generate_complex_gaussian_noise"]((3, 5, 4, 2), std=0.25, seed=42, xp=np)
    out["noise_1"] = sim["generate_complex_gaussian_noise"]((2, 7, 3), std=np.complex64(0.5 - 0.125j), seed=9, xp=np)
    # sample_disorder: (K0, K1, U0, U1, R, seed)
    dis_cases = [(8, 8, 4, 4, 1, 42), (16, 12, 4, 4, 2, 42), (12, 16, 4, 2, 4, 3)]
    for i, (k0, k1, u0, u1, r, sd) in enumerate(dis_cases):
        shot_ids, temporal, y_order, x_order = sim["sample_disorder"]((k0, k1), (u0, u1), R=r, seed=sd)
        out[f"dis_{i}_args"] = np.array([k0, k1, u0, u1, r, sd], dtype=np.int64)
        out[f"dis_{i}_shot_ids"] = shot_ids
        out[f"dis_{i}_temporal"] = temporal
        out[f"dis_{i}_y"] = y_order
        out[f"dis_{i}_x"] = x_order
    path = os.path.join(HERE, "motion_generators.npz")
    np.savez(path, **out)
    print(f"wrote {path}: {len(out)} arrays")
    disorder_vectors(ref)


# (K0, K1, tile0, tile1, accel, seed): the reference builds DisorderConfig from
# the image's (partition, line) extents, tile = (block_par, block_lin) and
# pat_lines_to_measure = K1 // accel (simulation.py:385-392)
DISORDER_CASES = [
    (64, 64, 4, 4, 1, 123456),      # 64^3, 16 shots
    (128, 128, 8, 8, 2, 123456),    # 128x128x96, 32 shots
    (192, 192, 8, 12, 2, 123456),   # 192x192x128, 48 shots
    (256, 256, 16, 16, 2, 123456),  # 256x256x192, 64 shots
    (96, 96, 12, 8, 2, 123456),     # a non-square tile of the batched sizes
    (48, 64, 4, 8, 2, 7),           # non-square plane and tile
    (30, 50, 4, 8, 3, 2**40 + 5),   # extents padded to tile multiples, 41-bit seed
]


def disorder_vectors(ref):
    import math
    import random
    import typing
    from dataclasses import dataclass

    sim = load_functions(os.path.join(ref, "simulation.py"), ["DisorderConfig", "make_disorder_order"],
                         extra={"random": random, "ceil": math.ceil, "dataclass": dataclass, "List": typing.List,
                                "Tuple": typing.Tuple})
    out = {}
    for i, (k0, k1, u0, u1, r, sd) in enumerate(DISORDER_CASES):
        cfg = sim["DisorderConfig"](max_line_number=k1 - 1, max_partition_number=k0 - 1,
                                    pat_lines_to_measure=k1 // r, block_lin=u1, block_par=u0, seed=sd)
        order, (L, P) = sim["make_disorder_order"](cfg)
        out[f"case_{i}_args"] = np.array([k0, k1, u0, u1, r, sd], dtype=np.int64)
        out[f"case_{i}_order"] = np.array(order, dtype=np.int16)
        out[f"case_{i}_LP"] = np.array([L, P], dtype=np.int32)
    path = os.path.join(HERE, "disorder_orders.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
