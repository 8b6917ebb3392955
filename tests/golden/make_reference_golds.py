"""Extract the MPFA-O regression golds and the IC vector that the reference's
own regression suite holds (demo/*/ *.regression.gold, the PETSc-binary IC
vector) into tests/golden/reference_golds.json.  Run in the development
container where the reference tree is mounted; the JSON it writes is data
(expected outputs plus the case definitions taken from the .cfg files)."""
import json
import os
import shutil
import sys

import numpy as np

REF = os.environ.get("TDYCORE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_gold(path):
    out = {"samples": {}}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("--"):
                continue
            k, v = line.split(":")
            k = k.strip()
            if k in ("Max", "Min", "Mean"):
                out[k] = float(v)
            else:
                out["samples"][k] = float(v)
    return out


# Case definitions restate demo/transient/transient_mpfaof90.{F90,cfg},
# demo/transient/transient_snes_mpfaof90.{F90,cfg}, demo/richards/richards.cfg
# and the option defaults of src/tdycore.c:260-296.
CASES = {
    "transient-mpfaof90-neumann": dict(
        gold="demo/transient/transient-mpfaof90-neumann.regression.gold",
        form="ts", mesh=[3, 3, 3], bc="neumann", gmatrix="mpfao", density="constant",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=100.0, nsteps=10, bc_fn="const1e5", nsample=2),
    "transient-mpfaof90-tpf-neumann": dict(
        gold="demo/transient/transient-mpfaof90-tpf-neumann.regression.gold",
        form="ts", mesh=[3, 3, 3], bc="neumann", gmatrix="tpf", density="constant",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=100.0, nsteps=10, bc_fn="const1e5", nsample=2),
    "transient-mpfaof90-tpf": dict(
        gold="demo/transient/transient-mpfaof90-tpf.regression.gold",
        form="ts", mesh=[3, 3, 3], bc="dirichlet", gmatrix="tpf", density="constant",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=100.0, nsteps=10, bc_fn="const1e5", nsample=2),
    "transient-mpfaof90": dict(
        gold="demo/transient/transient-mpfaof90.regression.gold",
        form="ts", mesh=[3, 3, 3], bc="dirichlet", gmatrix="mpfao", density="constant",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=100.0, nsteps=10, bc_fn="const1e5", nsample=2),
    "transient-snes-mpfaof90": dict(
        gold="demo/transient/transient-snes-mpfaof90.regression.gold",
        form="snes", mesh=[1, 1, 15], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K=1e-10, porosity=0.115, Sr=0.115, ic=102325.0, dt=1800.0, nsteps=10, bc_fn=None, nsample=2),
    "transient-snes-mpfaof90-tdydriver": dict(
        gold="demo/transient/transient-snes-mpfaof90-tdydriver.regression.gold",
        form="snes", mesh=[1, 1, 15], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K=1e-10, porosity=0.115, Sr=0.115, ic=102325.0, dt=1800.0, nsteps=10, bc_fn=None, nsample=2),
    "transient-snes-mpfaof90-dirichlet": dict(
        gold="demo/transient/transient-snes-mpfaof90-dirichlet.regression.gold",
        form="snes", mesh=[5, 5, 5], bc="dirichlet", gmatrix="tpf", density="exponential",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=1800.0, nsteps=48, bc_fn="top110k", nsample=2),
    "transient-snes-mpfaof90-seepage": dict(
        gold="demo/transient/transient-snes-mpfaof90-seepage.regression.gold",
        form="snes", mesh=[5, 5, 5], bc="seepage", gmatrix="tpf", density="exponential",
        K=1e-10, porosity=0.115, Sr=0.115, ic=90325.0, dt=1800.0, nsteps=48, bc_fn="top110k", nsample=2),
    # prisms read from the reference's Exodus mesh (copied beside this script as data)
    "transient-snes-mpfaof90-wedge": dict(
        gold="demo/transient/transient-snes-mpfaof90-wedge.regression.gold",
        form="snes", mesh_file="2x2x2_wedge_mesh.exo", bc="dirichlet", gmatrix="tpf", density="exponential",
        K=1e-10, porosity=0.115, Sr=0.115, ic=102325.0, dt=1800.0, nsteps=3, bc_fn="top110k", nsample=10),
    # a terrain-following transect of hexahedra read from Exodus (-tdy_read_mesh), no-flow boundaries
    "transient-snes-mpfaof90-transect-neumann": dict(
        gold="demo/transient/transient-snes-mpfaof90-transect-neumann.regression.gold",
        form="snes", mesh_file="transect_44x1x3_hex_uniform_dz_mesh.exo", bc="neumann", gmatrix="mpfao",
        density="exponential", K=1e-10, porosity=0.115, Sr=0.115, ic=57180.0, dt=1800.0, nsteps=10,
        snes_rtol=1e-30, bc_fn=None, nsample=2),
    # a quadrilateral surface mesh extruded into 3 layers of hexahedra (max_steps keeps the
    # demo's default of 2, transient_snes_mpfaof90.F90:223)
    "transient-snes-mpfaof90-dmplex-extrude": dict(
        gold="demo/transient/transient-snes-mpfaof90-dmplex-extrude.regression.gold",
        form="snes", mesh_file="3x3_quad_surface_mesh.exo", extrude_layers=3, bc="dirichlet", gmatrix="tpf",
        density="exponential", K=1e-10, porosity=0.115, Sr=0.115, ic=102325.0, dt=1800.0, nsteps=2,
        bc_fn="top110k", nsample=10),
    "richards-driver-snes-prob1": dict(
        gold="demo/richards/richards-driver-snes-prob1.regression.gold",
        form="tdyti", mesh=[2, 2, 2], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K=1e-12, porosity=0.25, Sr=0.15, ic="vec", dt_init=0.1, growth=1.5, dt_max=600.0,
        final_time=3.1536e3, bc_fn=None, nsample=2),
    "richards-driver-snes-checkpoint-write": dict(
        gold="demo/richards/richards-driver-snes-checkpoint-write.regression.gold",
        form="tdyti", mesh=[2, 2, 2], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K=1e-12, porosity=0.25, Sr=0.15, ic="vec", dt_init=0.1, growth=1.5, dt_max=600.0,
        final_time=3.1536e3, bc_fn=None, nsample=2),
    # initial state and materials from HDF5 (demo/richards/initial.h5, copied beside this script)
    "richards-driver-snes-initial-cond-isotropic-k": dict(
        gold="demo/richards/richards-driver-snes-initial-cond-isotropic-k.regression.gold",
        form="tdyti", mesh=[5, 4, 3], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K="h5diag3:Permeability", porosity=0.25, Sr=0.15, ic="h5:IC", h5_file="initial.h5",
        dt_init=0.1, growth=1.5, dt_max=600.0, final_time=3.1536e3, bc_fn=None, nsample=1),
    # the reference solved these Newton systems with GMRES + ILU(0) at KSP rtol 1e-5, and this case's
    # final state depends on that inexact solve at the level of SNES rtol 1e-8; it is compared at 2e-8
    "richards-driver-snes-initial-cond-anisotropic-k": dict(
        gold="demo/richards/richards-driver-snes-initial-cond-anisotropic-k.regression.gold",
        form="tdyti", mesh=[5, 4, 3], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K="h5xyz:Permeability", porosity="h5:Porosity", Sr=0.15, ic="h5:fields/IC", h5_file="initial.h5",
        dt_init=0.1, growth=1.5, dt_max=600.0, final_time=3.1536e3, bc_fn=None, nsample=1, rtol=2e-8),
    # restarts from the checkpoint that -checkpoint-write leaves at t = 3153.6 s (dataset fields/IC,
    # src/tdyio.c:340-359): the IC is that run's final state, and the integrator starts again from dt_init
    "richards-driver-snes-checkpoint-read": dict(
        gold="demo/richards/richards-driver-snes-checkpoint-read.regression.gold",
        form="tdyti", mesh=[2, 2, 2], bc="dirichlet", gmatrix="mpfao", density="exponential",
        K=1e-12, porosity=0.25, Sr=0.15, ic="vec", dt_init=0.1, growth=1.5, dt_max=600.0,
        final_time=3.1536e3, bc_fn=None, nsample=2, restart=1),
}


def main():
    out = {"cases": {}, "vm": 0.8, "valpha": 1e-4}
    for name, case in CASES.items():
        c = dict(case)
        c["expected"] = parse_gold(os.path.join(REF, c.pop("gold")))
        if "mesh_file" in c:
            shutil.copyfile(os.path.join(REF, "share/meshes", c["mesh_file"]), os.path.join(HERE, c["mesh_file"]))
        if "h5_file" in c:
            shutil.copyfile(os.path.join(REF, "demo/richards", c["h5_file"]), os.path.join(HERE, c["h5_file"]))
        out["cases"][name] = c
    raw = open(os.path.join(REF, "demo/richards/richards_driver_snes_prob1_3x3x3_random.vec"), "rb").read()
    hdr = np.frombuffer(raw[:8], dtype=">i4")
    assert int(hdr[0]) == 1211214
    out["ic_vec"] = [float(x) for x in np.frombuffer(raw[8:8 + 8 * int(hdr[1])], dtype=">f8")]
    with open(os.path.join(HERE, "reference_golds.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", len(out["cases"]), "cases")


if __name__ == "__main__":
    sys.exit(main())
