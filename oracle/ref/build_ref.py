#!/usr/bin/env python3
"""Build recipe for the reference GP-MOOD Fortran, used as a parity checker.

Test infrastructure only: nothing in the product imports or runs this.

The reference solver is compiled from its own sources where they lie under
``/root/reference/gp-mood-code`` (nothing is copied into this repository).
Outputs go only to ``oracle/_ref/<config>/`` (git-ignored).  What differs from
the reference's own Makefile (gp-mood-code/Makefile:8-58):

* compiler: ``amdflang`` (gfortran is not in this image) with the Makefile's
  ``-fdefault-real-8`` plus ``-O3 -ffp-contract=off``; ``-mcmodel=medium`` for
  grids of 2048 and more (the solution arrays are static module arrays);
* ``parameters.f90``: the reference takes its compile-time configuration from a
  ``module parameters`` copied out of ``problems/*.f90`` (README.md:10-12).  This
  script writes that module from the config table below, with the schema of
  ``problems/2DRP3_GPMOOD.f90:1-48``;
* ``main.f90``, ``output.f90`` and ``reader.f90`` (HDF5 I/O) are not built; the
  harness driver ``oracle/ref/ref_driver.f90`` replaces ``main.f90``;
* two compile errors under flang are resolved with preprocessor definitions,
  leaving the sources untouched:
  - ``mod_subroutine_MOOD.f90:197`` declares locals ``pR``/``pL`` that clash,
    case-insensitively, with the kind parameter ``PR``.  ``-DpR=pRgt -DpL=pLft``
    renames them (case-sensitive macro, so ``PR`` itself is untouched);
  - ``mod_NN.f90:41,44`` index ``values_NN`` with the NN input-size parameter
    ``L`` (the reference's undeclared loop variable ``l`` resolves to it), a
    compile-time bounds error when ``L`` exceeds the grid.  The NN arrays are
    never used with ``method = GP_MOOD``, so ``parameters.f90`` sets
    ``L = min(32, lf)``;
  - ``global_variables.f90:15`` evaluates ``szs_sphere(Mord+2)``, out of bounds
    for ``Mord=7`` (the value only feeds the reference NN stencil size).  The
    macro ``szs_sphere(x)=szs_sphere(min(x,7))`` clamps the index; flang's
    preprocessor only expands this function-like macro in ``-E`` mode, so this
    one file is preprocessed to a temporary directory inside the (git-ignored)
    output directory, compiled from there, and the temporary file is deleted.

The configurations marked ``cascade="full"`` build the reference with ONE line
changed, for the order cascade 7 -> 5 -> 3 -> 1 of the configured workloads:
the reference lowers a failing cell ``3 -> 1`` and ``>3 -> 3``
(``mod_subroutine_MOOD.f90:159-163``); the patched build replaces the statement
``CellGPO(l,n) = 3`` of the ``elseif (CellGPO(l,n) > 3)`` branch (line 162) by
``CellGPO(l,n) = CellGPO(l,n) - 2``.  Every other line of DETECTION, and every
other module, is the reference's own.  The patched copy is written to the same
temporary directory and deleted after compiling.

Usage: python oracle/ref/build_ref.py [config ...]   (default: all configs)
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

REF_SRC = "/root/reference/gp-mood-code"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_ROOT = os.path.join(os.path.dirname(HERE), "_ref")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

# Module build order of gp-mood-code/Makefile:27-58 without output/reader/main.
MODULES = [
    "constants", "parameters", "linalg", "global_variables", "GP_init",
    "mesh_init", "physics", "IC", "mod_append_NN_dataset", "mod_NN",
    "mod_HLLC", "mod_HLL", "mod_LLF", "mod_unsplit", "mod_GP_MOOD",
    "mod_NN_GP_MOOD_CC", "mod_GP_SE", "mod_FOG", "mod_pol_MOOD",
    "mod_Unlim_Pol", "mod_Reconstruction", "BC", "mod_subroutine_NN_MOOD",
    "mod_subroutine_MOOD", "mod_FE", "mod_SSP_RK2", "mod_SSP_RK3",
    "mod_SSP_RK4", "time_stepper",
]

BC_NAMES = {"neumann": "Neumann", "periodic": "Periodic", "reflective": "Reflective", "kh": "KH_BC",
            "dirichlet": "Dirichlet", "mach800": "Mach800_BC", "doublemach800": "DoubleMach800_BC", "rmi": "RMI_BC",
            "dmr": "DMR", "rt": "RT_BC"}
FLUX_NAMES = {"hllc": "HLLC", "llf": "LLF", "hll": "HLL", "unsplit": "unsplit"}
INTEGRATOR_NAMES = {"fe": "FE", "ssp_rk2": "SSP_RK2", "ssp_rk3": "SSP_RK3", "ssp_rk4": "SSP_RK4"}

# name: (N, Mord, ngp, bc, DMP, U2, U2_tol[, numFlux]); numFlux defaults to HLLC
CONFIGS = {
    "r1_n32_neu": (32, 3, 2, "neumann", True, True, True),
    "r2_n32_neu": (32, 5, 3, "neumann", True, True, True),
    "r2_n64_neu": (64, 5, 3, "neumann", True, True, True),
    "r2_n64_per": (64, 5, 3, "periodic", True, True, True),
    "r2_n64_ref": (64, 5, 3, "reflective", True, True, True),
    "r2_n64_nou2": (64, 5, 3, "neumann", True, False, False),
    "r3_n32_ref": (32, 7, 4, "reflective", True, True, True),
    "r3_n64_ref": (64, 7, 4, "reflective", True, True, True),
    "r3_n64_per": (64, 7, 4, "periodic", True, True, True),
    "r2_n256_neu": (256, 5, 3, "neumann", True, True, True),
    "r3_n2048_ref": (2048, 7, 4, "reflective", True, True, True),
    # the other Riemann solvers and the Kelvin-Helmholtz boundary of the
    # reference's problem files (problems/2DRP3_GPMOOD.f90, KH_GPMOOD.f90,
    # shu_osher_GPMOOD.f90, Mach800_GPMOOD.f90)
    "r2_n64_neu_uns": (64, 5, 3, "neumann", True, True, True, "unsplit"),
    "r2_n64_neu_hll": (64, 5, 3, "neumann", True, True, True, "hll"),
    "r1_n32_ref_uns": (32, 3, 2, "reflective", True, True, True, "unsplit"),
    "r3_n64_per_hll": (64, 7, 4, "periodic", True, True, True, "hll"),
    "r3_n64_kh_uns": (64, 7, 4, "kh", True, True, True, "unsplit"),
    "r2_n64_kh_hllc": (64, 5, 3, "kh", True, True, True),
    # formula boundary conditions on rectangular grids (problems/dmr.f90,
    # Mach800_GPMOOD.f90, RT.f90 at reduced resolution; BC.f90:60-88, 137-381)
    "r1_dmr_64x16": (64, 3, 2, "dmr", True, True, True),
    "r2_dmr_64x16": (64, 5, 3, "dmr", True, True, True),
    "r1_m800_32": (32, 3, 2, "mach800", True, True, True, "hll"),
    "r1_rt_16x48": (16, 3, 2, "rt", True, True, True, "unsplit"),
    "r2_rt_16x48": (16, 5, 3, "rt", True, True, True, "unsplit"),
    "r1_dm800_32": (32, 3, 2, "doublemach800", True, True, True, "hll"),
    "r1_rmi_48x8": (48, 3, 2, "rmi", True, True, True),
    "r1_dir_64x8": (64, 3, 2, "dirichlet", True, True, True, "llf"),
    # the other integrators (time_stepper.f90:17-21) and the LLF flux
    "r2_n32_neu_fe": (32, 5, 3, "neumann", True, True, True),
    "r2_n32_neu_rk2": (32, 5, 3, "neumann", True, True, True),
    "r2_n32_ref_rk4": (32, 5, 3, "reflective", True, True, True),
    "r3_n32_per_rk4": (32, 7, 4, "periodic", True, True, True, "hll"),
    "r2_n32_neu_llf": (32, 5, 3, "neumann", True, True, True, "llf"),
    "r1_dmr_64x16_rk2": (64, 3, 2, "dmr", True, True, True),
    "r1_m800_32_fe": (32, 3, 2, "mach800", True, True, True, "hll"),
    "r1_rt_16x48_rk2": (16, 3, 2, "rt", True, True, True, "unsplit"),
    # SSP-RK4 with formula boundaries and the RT source
    "r1_dmr_64x16_rk4": (64, 3, 2, "dmr", True, True, True),
    "r1_rt_16x48_rk4": (16, 3, 2, "rt", True, True, True, "unsplit"),
    "r1_m800_32_rk4": (32, 3, 2, "mach800", True, True, True, "hll"),
    # non-square cells (dx /= dy): the GP tables (ell from min(dx, dy),
    # GP_init.f90), Setdt's two directions and dt/dx, dt/dy of the update
    "r2_32x48_neu_rect": (32, 5, 3, "neumann", True, True, True),
    "r3_n32_ref_rect": (32, 7, 4, "reflective", True, True, True),
    # first-order Godunov (method = FOG, ngp = 1, FE) as the problems/*_FOG.f90 files
    "fog_n64_neu_uns": (64, 3, 1, "neumann", True, True, True, "unsplit"),
    "fog_n64_kh_uns": (64, 3, 1, "kh", True, True, True, "unsplit"),
    "fog_m800_32_hll": (32, 3, 1, "mach800", True, True, True, "hll"),
    "fog_n64_per_hll": (64, 3, 1, "periodic", True, True, True, "hll"),
    "fog_n32_ref_hllc": (32, 3, 1, "reflective", True, True, True),
    "fog_dmr_64x16": (64, 3, 1, "dmr", True, True, True),
    "fog_rt_16x48_uns": (16, 3, 1, "rt", True, True, True, "unsplit"),
    # the reference's own network (method NN_GP_MOOD / NN_GP_MOOD_CC; mod_NN.f90,
    # weights read by load_NN from refnn.txt in the working directory)
    "nn_r1_n64_neu": (64, 3, 2, "neumann", True, True, True),
    "nn_r1_n64_per_l52": (64, 3, 2, "periodic", True, True, True, "unsplit"),
    "nn_r1_64x48_kh_rk2": (64, 3, 2, "kh", True, True, True),
    "nn_r2_n128_ref": (128, 5, 3, "reflective", True, True, True),
    "nncc_r1_n64_neu": (64, 3, 2, "neumann", True, True, True),
    # polynomial MOOD (method = POL_MOOD, mod_pol_MOOD.f90; Mord 3, ngp 2)
    "pol_r1_n64_neu": (64, 3, 2, "neumann", True, True, True),
    "pol_r1_n32_ref_hll": (32, 3, 2, "reflective", True, True, True, "hll"),
    "pol_r1_dmr_64x16": (64, 3, 2, "dmr", True, True, True),
    "pol_r1_n64_per_nou2": (64, 3, 2, "periodic", True, False, False, "unsplit"),
    # one pass without detection: unlimited GP (GP_SE, mod_GP_SE.f90) and
    # unlimited polynomials (Unlim_POL, mod_Unlim_Pol.f90; Mord 3, ngp 2)
    "gpse_r2_n64_per": (64, 5, 3, "periodic", True, True, True),
    "gpse_r3_n32_neu_fe": (32, 7, 4, "neumann", True, True, True, "unsplit"),
    "gpse_r2_rt_16x48": (16, 5, 3, "rt", True, True, True, "unsplit"),
    "unlim_r1_n64_kh_rk2": (64, 3, 2, "kh", True, True, True, "hll"),
    # FOG with the multi-stage integrators
    "fog_n64_per_rk3": (64, 3, 1, "periodic", True, True, True),
    "fog_dmr_64x16_rk2": (64, 3, 1, "dmr", True, True, True, "hll"),
    # the order cascade 7 -> 5 -> 3 -> 1 (one-line patched DETECTION, see the
    # module docstring) at GP-R3 under the three copy boundaries
    "r3_n64_ref_full": (64, 7, 4, "reflective", True, True, True),
    "r3_n64_per_full": (64, 7, 4, "periodic", True, True, True),
    "r3_n64_neu_full": (64, 7, 4, "neumann", True, True, True),
    "r3_n32_ref_full": (32, 7, 4, "reflective", True, True, True),
    "r3_n2048_ref_full": (2048, 7, 4, "reflective", True, True, True),
    # write_NN_dataset = .true.: the training samples of append_to_NN_datatset
    # (mod_append_NN_dataset.f90:11-79), dumped by the harness driver
    "dset_r1_n64_neu": (64, 3, 2, "neumann", True, True, True),
}

# The problem files' own configurations (problems/*.f90: grid, domain, method,
# integrator, numFlux, BC_type, Mord, ngp, DMP/U2 switches), for checksum runs.
PROBLEM_FILES = {
    # name: (problem file, lf, nf, Lx, Ly, method, integrator, numFlux, BC, Mord, ngp, DMP, problem)
    "pf_2drp3": ("2DRP3_GPMOOD.f90", 256, 256, "1.", "1.", "GP_MOOD", "ssp_rk3", "unsplit", "neumann", 3, 2, True, "RP_2D_3"),
    "pf_2drp3_nodmp": ("2DRP3_GPMOOD_noDMP.f90", 256, 256, "1.", "1.", "GP_MOOD", "ssp_rk3", "unsplit", "neumann", 3, 2,
                       False, "RP_2D_3"),
    "pf_2drp12": ("2DRP12.f90", 256, 256, "1.", "1.", "GP_MOOD", "ssp_rk3", "hllc", "neumann", 3, 2, True, "RP_2D_12"),
    "pf_kh": ("KH_GPMOOD.f90", 256, 256, "1.", "1.", "GP_MOOD", "ssp_rk3", "unsplit", "kh", 3, 2, True, "KH"),
    "pf_mach800": ("Mach800_GPMOOD.f90", 256, 256, "1.5", "1.5", "GP_MOOD", "ssp_rk3", "hll", "mach800", 3, 2, True,
                   "Mach800"),
    "pf_mach800_nodmp": ("Mach800_GPMOOD_noDMP.f90", 256, 256, "1.5", "1.5", "GP_MOOD", "ssp_rk3", "hll", "mach800", 3,
                         2, False, "Mach800"),
    "pf_rt": ("RT.f90", 64, 192, "0.5", "1.5", "GP_MOOD", "ssp_rk3", "unsplit", "rt", 3, 2, True, "RT"),
    "pf_dmr": ("dmr.f90", 512, 128, "4.", "1.", "GP_MOOD", "ssp_rk3", "hllc", "dmr", 3, 2, True, "DMR"),
    "pf_explosion": ("explosion.f90", 256, 256, "1.5", "1.5", "GP_MOOD", "ssp_rk3", "hllc", "neumann", 3, 2, True,
                     "explosion"),
    "pf_implosion": ("implosion.f90", 256, 256, "0.3", "0.3", "GP_MOOD", "ssp_rk3", "hllc", "reflective", 3, 2, True,
                     "implosion"),
    "pf_shu_osher": ("shu_osher_GPMOOD.f90", 512, 512, "28.2842712474619", "28.2842712474619", "GP_MOOD", "ssp_rk3",
                     "hll", "periodic", 3, 2, True, "Shu_Osher_rotated"),
    "pf_2drp3_fog": ("2DRP3_FOG.f90", 256, 256, "1.", "1.", "FOG", "fe", "unsplit", "neumann", 3, 1, True, "RP_2D_3"),
    "pf_kh_fog": ("KH_FOG.f90", 256, 256, "1.", "1.", "FOG", "fe", "unsplit", "kh", 3, 1, True, "KH"),
    "pf_mach800_fog": ("Mach800_FOG.f90", 256, 256, "1.5", "1.5", "FOG", "fe", "hll", "mach800", 3, 1, True, "Mach800"),
    "pf_shu_osher_fog": ("shu_osher_FOG.f90", 512, 512, "28.2842712474619", "28.2842712474619", "FOG", "fe", "hll",
                         "periodic", 3, 1, True, "Shu_Osher_rotated"),
    "pf_sedov_fog": ("sedov_FOG.f90", 256, 256, "1.", "1.", "FOG", "fe", "unsplit", "neumann", 3, 1, True, "sedov"),
}
for _k, (_f, _lf, _nf, _lx, _ly, _m, _i, _fl, _bc, _mo, _ng, _dmp, _pb) in PROBLEM_FILES.items():
    CONFIGS[_k] = (_lf, _mo, _ng, _bc, _dmp, _dmp, _dmp, _fl)  # the noDMP files also switch U2, U2_tol off

# grid, domain, integrator and problem of the configurations that differ from
# a square unit domain, SSP_RK3 and a problem without source term
EXTRAS = {
    "r1_dmr_64x16": dict(nf=16, lx="4.", ly="1."),
    "r2_dmr_64x16": dict(nf=16, lx="4.", ly="1."),
    "r1_m800_32": dict(lx="1.5", ly="1.5"),
    "r1_rt_16x48": dict(nf=48, lx="0.5", ly="1.5", problem="RT"),
    "r2_rt_16x48": dict(nf=48, lx="0.5", ly="1.5", problem="RT"),
    "r1_dm800_32": dict(lx="1.5", ly="1.5"),
    "r1_rmi_48x8": dict(nf=8, lx="6.", ly="1."),
    "r1_dir_64x8": dict(nf=8, lx="9.", ly="1.125"),
    "r2_n32_neu_fe": dict(integrator="fe"),
    "r2_n32_neu_rk2": dict(integrator="ssp_rk2"),
    "r2_n32_ref_rk4": dict(integrator="ssp_rk4"),
    "r3_n32_per_rk4": dict(integrator="ssp_rk4"),
    "r1_dmr_64x16_rk2": dict(nf=16, lx="4.", ly="1.", integrator="ssp_rk2"),
    "r1_m800_32_fe": dict(lx="1.5", ly="1.5", integrator="fe"),
    "r1_rt_16x48_rk2": dict(nf=48, lx="0.5", ly="1.5", problem="RT", integrator="ssp_rk2"),
    "r1_dmr_64x16_rk4": dict(nf=16, lx="4.", ly="1.", integrator="ssp_rk4"),
    "r1_rt_16x48_rk4": dict(nf=48, lx="0.5", ly="1.5", problem="RT", integrator="ssp_rk4"),
    "r1_m800_32_rk4": dict(lx="1.5", ly="1.5", integrator="ssp_rk4"),
    "r2_32x48_neu_rect": dict(nf=48, lx="1.", ly="1."),
    "r3_n32_ref_rect": dict(lx="2.", ly="1."),
    "fog_n64_neu_uns": dict(integrator="fe", method="FOG"),
    "fog_n64_kh_uns": dict(integrator="fe", method="FOG"),
    "fog_m800_32_hll": dict(lx="1.5", ly="1.5", integrator="fe", method="FOG"),
    "fog_n64_per_hll": dict(integrator="fe", method="FOG"),
    "fog_n32_ref_hllc": dict(integrator="fe", method="FOG"),
    "fog_dmr_64x16": dict(nf=16, lx="4.", ly="1.", integrator="fe", method="FOG"),
    "fog_rt_16x48_uns": dict(nf=48, lx="0.5", ly="1.5", problem="RT", integrator="fe", method="FOG"),
    "nn_r1_n64_neu": dict(method="NN_GP_MOOD", nn_l=57, nn_file="refnn"),
    "nn_r1_n64_per_l52": dict(method="NN_GP_MOOD", nn_l=52, nn_file="refnn"),
    "nn_r1_64x48_kh_rk2": dict(nf=48, method="NN_GP_MOOD", integrator="ssp_rk2", nn_l=57, nn_file="refnn"),
    "nn_r2_n128_ref": dict(method="NN_GP_MOOD", nn_l=100, nn_len=48, nn_file="refnn"),
    "nncc_r1_n64_neu": dict(method="NN_GP_MOOD_CC", nn_l=57, nn_file="refnn"),
    "pol_r1_n64_neu": dict(method="POL_MOOD"),
    "pol_r1_n32_ref_hll": dict(method="POL_MOOD", integrator="ssp_rk2"),
    "pol_r1_dmr_64x16": dict(nf=16, lx="4.", ly="1.", method="POL_MOOD"),
    "pol_r1_n64_per_nou2": dict(method="POL_MOOD"),
    "gpse_r2_n64_per": dict(method="GP_SE"),
    "gpse_r3_n32_neu_fe": dict(method="GP_SE", integrator="fe"),
    "gpse_r2_rt_16x48": dict(nf=48, lx="0.5", ly="1.5", problem="RT", method="GP_SE"),
    "unlim_r1_n64_kh_rk2": dict(method="Unlim_POL", integrator="ssp_rk2"),
    "fog_n64_per_rk3": dict(method="FOG"),
    "fog_dmr_64x16_rk2": dict(nf=16, lx="4.", ly="1.", method="FOG", integrator="ssp_rk2"),
    "r3_n64_ref_full": dict(cascade="full"),
    "r3_n64_per_full": dict(cascade="full"),
    "r3_n64_neu_full": dict(cascade="full"),
    "r3_n32_ref_full": dict(cascade="full"),
    "r3_n2048_ref_full": dict(cascade="full"),
    "dset_r1_n64_neu": dict(write_nn=True, dataset_size=60000, nn_l=52),
}

# DETECTION's cascade branch (mod_subroutine_MOOD.f90:161-162) and its
# replacement in the cascade="full" builds
CASCADE_LINE = ("               elseif (CellGPO(l,n) >  3)  then\n"
                "                  CellGPO(l,n) = 3\n")
CASCADE_FULL = ("               elseif (CellGPO(l,n) >  3)  then\n"
                "                  CellGPO(l,n) = CellGPO(l,n) - 2\n")


def patched_mood_source(src, dst):
    """Write mod_subroutine_MOOD.f90 with the one-line cascade change to dst."""
    text = open(src).read()
    if text.count(CASCADE_LINE) != 1:
        raise RuntimeError("cascade branch of mod_subroutine_MOOD.f90 not found exactly once")
    with open(dst, "w") as f:
        f.write(text.replace(CASCADE_LINE, CASCADE_FULL))


def extras(name):
    """nf, Lx, Ly (as Fortran literals), integrator, problem, method and
    cascade of a configuration."""
    e = dict(nf=CONFIGS[name][0], lx="1.", ly="1.", integrator="ssp_rk3", problem="RP_2D_3", method="GP_MOOD",
             cascade="reference")
    e.update(EXTRAS.get(name, {}))
    if name in PROBLEM_FILES:
        _, _, nf, lx, ly, method, integ, _, _, _, _, _, problem = PROBLEM_FILES[name]
        e.update(nf=nf, lx=lx, ly=ly, method=method, integrator=integ, problem=problem)
    return e


def config(name):
    """(N, Mord, ngp, bc, DMP, U2, U2_tol, flux) of a configuration."""
    c = CONFIGS[name]
    return tuple(c[:7]) + ((c[7] if len(c) > 7 else "hllc"),)

# configs for which the Fortran drop-in (reference modules + product
# time_stepper over libgpmood) is also built
DROPIN_CONFIGS = ["r2_n64_neu", "r3_n64_ref", "r3_n64_kh_uns", "r1_dmr_64x16", "r1_rt_16x48", "r1_m800_32_fe",
                  "r2_n32_ref_rk4", "fog_n64_kh_uns", "fog_m800_32_hll", "nn_r1_n64_neu", "nncc_r1_n64_neu",
                  "pol_r1_n64_neu", "r3_n32_ref_rect", "r1_rt_16x48_rk4", "dset_r1_n64_neu"]

PARAMS_TEMPLATE = """module parameters
   use constants
   implicit none
   real(PR), parameter :: CFL  =  0.8
   integer , parameter :: integrator    = {integrator}
   logical , parameter :: restart = .false.
   character(100) :: restart_filename='.'
   integer, parameter :: method   = {method}
   real(16), parameter :: ell_o_dx = 12.0
   integer, parameter :: Mord= {mord}
   integer, parameter :: ngp = {ngp}
   logical , parameter :: DMP        = {dmp}
   logical , parameter :: U2         = {u2}
   logical , parameter :: U2_tol     = {u2tol}
   logical :: write_NN_dataset={write_nn}
   integer, parameter :: dataset_size = {dataset_size}
   integer, parameter :: L={nn_l}
   integer, parameter :: length={nn_len}
   character(100) :: NN_filename='{nn_file}'
   integer, parameter :: numFlux = {flux}
   integer, parameter :: IO_freqStep = -100
   real(PR), parameter:: IO_freqTime = -1.e-3
   integer,  parameter :: lf = {n}
   integer,  parameter :: nf = {nf}
   integer, parameter  :: problem = {problem}
   real(PR), parameter :: tmax = 1.e10
   integer, parameter  :: nmax = 100000
   real(16), parameter :: Lx_16 = {lx}
   real(16), parameter :: Ly_16 = {ly}
   integer, parameter  :: BC_type = {bc}
end module
"""


def fbool(b):
    return ".true." if b else ".false."


def build(name, dropin=False):
    """Build the reference solver for config `name`.  With dropin=True the
    reference's time_stepper.f90 is replaced by the product's Fortran drop-in
    (gp-mood_amd/fortran/time_stepper.f90 over gpm_capi.f90), which runs the
    step through libgpmood; everything else is the reference's own code."""
    n, mord, ngp, bc, dmp, u2, u2tol, flux = config(name)
    e = extras(name)
    out = os.path.join(OUT_ROOT, name + ("_dropin" if dropin else ""))
    os.makedirs(out, exist_ok=True)
    params = os.path.join(out, "parameters.f90")
    with open(params, "w") as f:
        f.write(PARAMS_TEMPLATE.format(mord=mord, ngp=ngp, dmp=fbool(dmp), u2=fbool(u2),
                                       u2tol=fbool(u2tol), n=n, bc=BC_NAMES[bc], flux=FLUX_NAMES[flux],
                                       nf=e["nf"], lx=e["lx"], ly=e["ly"], problem=e["problem"],
                                       integrator=INTEGRATOR_NAMES[e["integrator"]], nn_l=e.get("nn_l", min(32, n)),
                                       nn_len=e.get("nn_len", 60), nn_file=e.get("nn_file", ""),
                                       method=e["method"], write_nn=fbool(e.get("write_nn", False)),
                                       dataset_size=e.get("dataset_size", 1)))
    flags = ["-fdefault-real-8", "-O3", "-ffp-contract=off", "-J", out, "-I", out]
    if n * extras(name)["nf"] >= 1024 * 1024:
        flags.append("-mcmodel=medium")
    objs = []
    tmpdir = tempfile.mkdtemp(prefix="pp_", dir=out)
    try:
        shim = os.path.join(os.path.dirname(OUT_ROOT), "..", "gp-mood_amd", "fortran")
        for mod in MODULES + (["__shim__"] if dropin else []):
            if dropin and mod == "time_stepper":
                continue
            if mod == "__shim__":
                for f in ("gpm_capi", "time_stepper"):
                    obj = os.path.join(out, f + "_shim.o")
                    subprocess.run([FC] + flags + ["-c", os.path.join(shim, f + ".f90"), "-o", obj], check=True, cwd=out)
                    objs.append(obj)
                continue
            src = params if mod == "parameters" else os.path.join(REF_SRC, mod + ".f90")
            obj = os.path.join(out, mod + ".o")
            extra = []
            if mod == "mod_subroutine_MOOD":
                extra = ["-cpp", "-DpR=pRgt", "-DpL=pLft"]
                if e["cascade"] == "full":
                    pp = os.path.join(tmpdir, "mod_subroutine_MOOD.f90")
                    patched_mood_source(src, pp)
                    src = pp
            if mod == "global_variables" and mord == 7:
                pp = os.path.join(tmpdir, "global_variables_pp.f90")
                with open(pp, "w") as f:
                    subprocess.run([FC, "-cpp", "-E", "-P", "-Dszs_sphere(x)=szs_sphere(min(x,7))", src],
                                   check=True, stdout=f)
                src = pp
            subprocess.run([FC] + flags + extra + ["-c", src, "-o", obj], check=True, cwd=out)
            objs.append(obj)
        drv = os.path.join(HERE, "ref_driver.f90")
        dobj = os.path.join(out, "ref_driver.o")
        subprocess.run([FC] + flags + ["-c", drv, "-o", dobj], check=True, cwd=out)
        exe = os.path.join(out, "gpm_dropin.exe" if dropin else "gpm_ref.exe")
        link = [FC] + [f for f in flags if f not in ("-J", "-I", out)] + objs + [dobj, "-o", exe]
        if dropin:
            libdir = os.path.abspath(os.path.join(shim, ".."))
            link += ["-L" + libdir, "-lgpmood", "-Wl,-rpath,$ORIGIN/../../../gp-mood_amd"]
        subprocess.run(link, check=True, cwd=out)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    for f in os.listdir(out):
        if f.endswith(".o") or f.endswith(".mod"):
            os.remove(os.path.join(out, f))
    return exe


def main(argv):
    if not os.path.isdir(REF_SRC):
        print("reference sources not present; nothing to build", file=sys.stderr)
        return 0
    names = [a for a in argv if not a.startswith("--")] or list(CONFIGS)
    jobs = [(nm, False) for nm in names]
    if "--no-dropin" not in argv:
        jobs += [(nm, True) for nm in DROPIN_CONFIGS if nm in names]
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for exe in ex.map(lambda j: build(*j), jobs):
            print("built", exe)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
