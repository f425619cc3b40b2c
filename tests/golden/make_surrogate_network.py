"""Train the surrogate network that tests/test_gpu_surrogate_trees.py loads
(needs a GPU): the training recipe of bench.py's surrogate leg for one seed --
16e6 intervals of non-satellite halos from merger-tree trajectories (forests
of 1024 trees drawn as the 64-tree workload: roots log-uniform 1e10-1e12 Msun,
resolution 1e9 Msun, z=5 -> 0, forest seeds 100, 101, ...), 30 epochs at
batch 8192, network seed 1.  Writes surrogate_tree_network.npz beside this
file (or to the path given): the float32 parameters (the network's weights
run as bf16 and its biases as float32) and the input/output normalisation.

    python tests/golden/make_surrogate_network.py [out.npz]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main(out):
    import torch
    from galacticus_amd import params as gp, surrogate as sg
    from galacticus_amd.cooling import synthetic_cie_table
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    table = synthetic_cie_table()
    pairs = sg.generate_tree_pairs(physics, table, 16_000_000, seed=100)
    ev = NodeEvolverBatched(1024, physics, table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    mlp, norm, rep = sg.train_surrogate(ev, pairs, epochs=30, batch=8192, seed=1)
    print(f"trained {rep.steps} steps in {rep.seconds:.1f} s, loss {rep.loss_first:.4g} -> {rep.loss_last:.4g}",
          flush=True)
    np.savez_compressed(out, params=mlp.get_params().astype(np.float32), norm=norm.cpu().numpy().astype(np.float32))
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "surrogate_tree_network.npz"))
