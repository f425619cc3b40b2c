"""Galacticus-layout output of evolved forests.

The reference writes, per output time, ``Outputs/OutputN/nodeData/<property>``
datasets (one entry per node that exists at that time, the nodes of each tree
contiguous), with ``Outputs/OutputN/mergerTreeIndex``,
``mergerTreeStartIndex`` and ``mergerTreeCount`` locating each tree's block
(merger_trees/outputter/standard.F90:235-392; the layout its tests read, e.g.
testSuite/test-output-tree-contiguousity.py and test-reproducibility.py).
This module writes the same group/dataset layout into a NumPy ``.npz``
archive (HDF5 is not a dependency of this package): the keys are the HDF5
paths, ``read_output`` returns them as nested dictionaries, so an analysis
that indexes ``f["Outputs/Output1/nodeData/diskMassStellar"]`` runs
unchanged on the dictionary.
"""
from __future__ import annotations

import numpy as np

from . import workloads as wl

# Galacticus output names of the state variables this engine evolves
# (component property names of the standard hot halo, disk, spheroid and
# black hole; "Metals" is the abundance of the single metal tracer).
OUTPUT_PROPERTIES = {
    "hotHaloMass": "HOT_MASS",
    "hotHaloAbundancesMetals": "HOT_ABUNDANCES",
    "hotHaloAngularMomentum": "HOT_ANGULAR_MOMENTUM",
    "hotHaloOuterRadius": "HOT_OUTER_RADIUS",
    "hotHaloOutflowedMass": "HOT_OUTFLOWED_MASS",
    "hotHaloOutflowedAbundancesMetals": "HOT_OUTFLOWED_ABUNDANCES",
    "hotHaloOutflowedAngularMomentum": "HOT_OUTFLOWED_ANGULAR_MOMENTUM",
    "hotHaloUnaccretedMass": "HOT_UNACCRETED_MASS",
    "hotHaloUnaccretedAbundancesMetals": "HOT_UNACCRETED_ABUNDANCES",
    "diskMassGas": "DISK_MASS_GAS",
    "diskMassStellar": "DISK_MASS_STELLAR",
    "diskAbundancesGasMetals": "DISK_ABUNDANCES_GAS",
    "diskAbundancesStellarMetals": "DISK_ABUNDANCES_STELLAR",
    "diskAngularMomentum": "DISK_ANGULAR_MOMENTUM",
    "spheroidMassGas": "SPHEROID_MASS_GAS",
    "spheroidMassStellar": "SPHEROID_MASS_STELLAR",
    "spheroidAbundancesGasMetals": "SPHEROID_ABUNDANCES_GAS",
    "spheroidAbundancesStellarMetals": "SPHEROID_ABUNDANCES_STELLAR",
    "spheroidAngularMomentum": "SPHEROID_ANGULAR_MOMENTUM",
    "blackHoleMass": "BH_MASS",
    "blackHoleSpin": "BH_SPIN",
}


def surviving_nodes(forest, schedule) -> np.ndarray:
    """Nodes that exist after the schedule has run: every node except those
    whose galaxy was promoted into their parent (the child node is removed,
    node_evolver/standard.F90:1241-1356) and satellites whose galaxy merged
    into a host.  Sorted by node index, so the nodes of a tree are contiguous."""
    from .trees import OP_PROMOTE, OP_MERGE_GALAXY
    gone = np.zeros(forest.n_nodes, dtype=bool)
    ops = schedule.ops
    gone[ops["a"][(ops["kind"] == OP_PROMOTE) | (ops["kind"] == OP_MERGE_GALAXY)]] = True
    return np.nonzero(~gone)[0]


def satellites(schedule, n_nodes) -> np.ndarray:
    """True for nodes that became satellites during the schedule."""
    from .trees import OP_BECOME_SATELLITE
    sat = np.zeros(n_nodes, dtype=bool)
    ops = schedule.ops
    sat[ops["a"][ops["kind"] == OP_BECOME_SATELLITE]] = True
    return sat


def output_group(forest, schedule, y, output_time, expansion_factor=None) -> dict:
    """The contents of one Outputs/OutputN group for state y (numpy (30, n)):
    keys relative to the group."""
    nodes = surviving_nodes(forest, schedule)
    tree = forest.tree_of_node()[nodes]
    sat = satellites(schedule, forest.n_nodes)[nodes]
    group = {
        "nodeData/nodeIndex": nodes.astype(np.int64),
        "nodeData/parentIndex": forest.parent[nodes].astype(np.int64),
        "nodeData/nodeIsIsolated": (~sat).astype(np.int32),
        "nodeData/mergerTreeIndex": tree.astype(np.int64),
        "nodeData/basicMass": forest.mass[nodes].astype(np.float64),
    }
    for name, var in OUTPUT_PROPERTIES.items():
        group["nodeData/" + name] = np.ascontiguousarray(y[wl.V[var], nodes], dtype=np.float64)
    trees, start, count = np.unique(tree, return_index=True, return_counts=True)
    group["mergerTreeIndex"] = trees.astype(np.int64)
    group["mergerTreeStartIndex"] = start.astype(np.int64)
    group["mergerTreeCount"] = count.astype(np.int64)
    group["outputTime"] = np.float64(output_time)
    if expansion_factor is not None:
        group["outputExpansionFactor"] = np.float64(expansion_factor)
    return group


def write_output(path, groups: list) -> None:
    """Write output groups (from output_group, in output order) as
    Outputs/Output1, Outputs/Output2, ... into an .npz archive."""
    arrays = {}
    for i, group in enumerate(groups, start=1):
        for key, value in group.items():
            arrays[f"Outputs/Output{i}/{key}"] = np.asarray(value)
    np.savez_compressed(path, **arrays)


def read_output(path) -> dict:
    """Read an archive written by write_output: a dictionary keyed by the
    full paths ("Outputs/Output1/nodeData/diskMassStellar") and, for
    convenience, nested dictionaries under "Outputs"."""
    with np.load(path, allow_pickle=False) as f:
        flat = {k: f[k] for k in f.files}
    nested: dict = {}
    for key, value in flat.items():
        d = nested
        parts = key.split("/")
        for p in parts[:-1]:
            d = d.setdefault(p, {})
        d[parts[-1]] = value
    flat.update(nested)
    return flat
