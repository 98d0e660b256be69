"""Input decks of the reference: singleideal.ini.

``read_deck`` parses the INI format of the reference's configuration files:
``[section]`` headers, ``key = value`` lines and ``;`` comments.  Values keep
their text; a scalar is read as Fortran list-directed input reads it (the
first item before a comma or blank, ``D`` exponents allowed), and
``cp_coeff`` is a comma list.  ``apply_patch`` applies a ``tests/tests.ini``
patch string ("section,key=value;section,key=value") with the reference
runner's split rules (tests/run_all_tests.py ``parse_config_block``: each
pair split at its first comma and first '=').

``case_from_deck`` turns a channel deck (grid_type 3, flow_init 0) into a
ChannelCase with the keys, defaults and checks of the equation module's
input routine (singleideal.F90:304-860): the channel grid (generate_grid_cha),
the scalars of set_fluid_prop / set_flow_params / set_chan_prop, and the
initial field of init_channel with the deck's MPI splits for its random
draws.  ``run_deck`` runs the deck's num_iter steps on the GPU and writes
progress.out rows every print_control steps and the restart files
rst.bin / field_info0.dat at the end, as the reference's run does
(singleideal_cpu.F90:1599-1693).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass

from .cases import ChannelCase, init_channel_field
from .grid import channel_grid
from .params import FlowCase, channel_params


def read_deck(path) -> dict:
    """{section: {key: text}} of an INI deck (keys keep their case)."""
    deck: dict = {}
    sec = None
    with open(path) as f:
        for line in f:
            text = line.split(";", 1)[0].strip()
            if not text:
                continue
            m = re.match(r"^\[([^\]]+)\]$", text)
            if m:
                sec = m.group(1).strip()
                deck.setdefault(sec, {})
                continue
            if "=" not in text or sec is None:
                raise ValueError(f"{path}: cannot read line {line.rstrip()!r}")
            key, val = text.split("=", 1)
            deck[sec][key.strip()] = val.strip()
    return deck


def apply_patch(deck: dict, patch: str) -> dict:
    """A copy of ``deck`` with a tests.ini patch applied."""
    out = {s: dict(kv) for s, kv in deck.items()}
    for pair in patch.split(";"):
        if not pair.strip():
            continue
        section, rest = pair.split(",", 1)
        key, value = rest.split("=", 1)
        out.setdefault(section.strip(), {})[key.strip()] = value.strip()
    return out


def _item(text: str) -> str:
    t = text.strip()
    return re.split(r"[,\s]", t, maxsplit=1)[0]


def _real(text: str) -> float:
    return float(_item(text).replace("D", "E").replace("d", "e"))


def _int(text: str) -> int:
    return int(_item(text))


def _reals(text: str) -> tuple:
    return tuple(float(v.strip().replace("D", "E").replace("d", "e")) for v in text.split(",") if v.strip())


class _Keys:
    """Typed access with the reference's required / optional keys."""

    def __init__(self, deck: dict):
        self.deck = deck

    def has(self, sec, key):
        return key in self.deck.get(sec, {})

    def get(self, sec, key, conv, default=None):
        if not self.has(sec, key):
            if default is None:
                raise ValueError(f"deck: missing key {key} in [{sec}]")
            return default
        return conv(self.deck[sec][key])


@dataclass
class DeckRun:
    """A channel deck made ready to run: the case and the run controls."""
    case: ChannelCase
    num_iter: int
    print_control: int
    splits: tuple


def case_from_deck(deck, nx=None, ny=None, nz=None, precision: int = 0) -> DeckRun:
    """ChannelCase of a channel deck (a path or a read_deck dict); nx/ny/nz
    override nxmax/nymax/nzmax.  Decks this path does not cover (curvilinear
    grids, grid types other than CHA, flow_init other than the channel,
    rk_type other than Wray, rfac_type 1) raise ValueError."""
    if not isinstance(deck, dict):
        deck = read_deck(deck)
    k = _Keys(deck)
    if k.get("grid", "grid_dim", _int, 1) != 1:
        raise ValueError("deck: curvilinear grids (grid_dim 2) are not covered")
    if k.get("grid", "grid_type", _int) != 3:
        raise ValueError("deck: only the channel grid (grid_type 3) is generated here")
    if k.get("flow", "flow_init", _int) != 0:
        raise ValueError("deck: only the channel flow (flow_init 0) is set up here")
    if k.get("numerics", "rk_type", _int) != 1:
        raise ValueError("deck: rk_type must be 1 (Wray)")
    if k.get("fluid", "rfac_type", _int, 0) != 0:
        raise ValueError("deck: rfac_type 1 is not covered")
    ep_order = k.get("numerics", "ep_order", _int)
    if k.get("numerics", "visc_order", _int) != ep_order:
        raise ValueError("deck: viscous order must be equal to ep_order")
    nxmax = k.get("grid", "nxmax", _int) if nx is None else nx
    nymax = k.get("grid", "nymax", _int) if ny is None else ny
    nzmax = k.get("grid", "nzmax", _int) if nz is None else nz
    splits = (k.get("mpi", "x_split", _int), k.get("mpi", "y_split", _int), k.get("mpi", "z_split", _int))
    bcn = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
    cp = k.get("fluid", "calorically_perfect", _int)
    fc = FlowCase(
        nx=nxmax, ny=nymax, nz=nzmax, ng=k.get("grid", "ng", _int),
        lx=k.get("grid", "domain_size_x", _real), ly=k.get("grid", "domain_size_y", _real),
        lz=k.get("grid", "domain_size_z", _real),
        mach=k.get("flow", "Mach", _real), reynolds=k.get("flow", "Reynolds", _real),
        theta_wall=k.get("flow", "theta_wall", _real), gam=k.get("fluid", "gam", _real),
        prandtl=k.get("fluid", "Prandtl", _real), visc_model=k.get("fluid", "visc_model", _int),
        calorically_perfect=cp,
        indx_cp_l=k.get("fluid", "indx_cp_l", _int) if cp == 0 else 0,
        indx_cp_r=k.get("fluid", "indx_cp_r", _int) if cp == 0 else 0,
        cp_coeff=k.get("fluid", "cp_coeff", _reals) if cp == 0 else (),
        flux_splitting=k.get("numerics", "flux_splitting", _int, 0),
        conservative_viscous=k.get("numerics", "conservative_viscous", _int),
        vt_exp=k.get("fluid", "vt_exp", _real), s_suth=k.get("fluid", "s_suth", _real),
        t_ref=k.get("flow", "T_ref", _real), ep_order=ep_order,
        nkeep=k.get("numerics", "nkeep", _int, 0), weno_scheme=k.get("numerics", "weno_scheme", _int),
        weno_version=k.get("numerics", "weno_version", _int),
        sensor_threshold=k.get("numerics", "sensor_threshold", _real),
        cfl=k.get("controls", "cfl", _real),
        iter_dt_recompute=k.get("controls", "iter_dt_recompute", _int),
        bctags=tuple(k.get("bc", b, _int) for b in bcn),
        bctags_nr=tuple(k.get("bc", b + "_nr", _int) for b in bcn),
        correct_bound_ord=k.get("bc", "correct_bound_ord", _int, 0),
        dyptarget=k.get("grid", "dyptarget", _real), jbgrid=k.get("grid", "jbgrid", _real),
        ysmoosteps=k.get("grid", "ysmoosteps", _int), precision=precision)
    if any(n % s for n, s in zip((nxmax, nymax, nzmax), splits)):
        raise ValueError("deck: the grid must be divisible by the MPI splits")
    grid = channel_grid(nxmax, nymax, nzmax, ng=fc.ng, lx=fc.lx, ly=fc.ly, lz=fc.lz, retau=fc.reynolds,
                        jbgrid=fc.jbgrid, dyptarget=fc.dyptarget, nsmooy=fc.ysmoosteps,
                        metrics_order=k.get("grid", "metrics_order", _int))
    params = channel_params(fc, grid.yn)
    params.rand_type = k.get("numerics", "rand_type", _int)
    w0 = init_channel_field(grid, params, fc.lz, splits=splits, rand_type=params.rand_type)
    return DeckRun(ChannelCase(fc, grid, params, w0), k.get("controls", "num_iter", _int),
                   k.get("output", "print_control", _int, 1), splits)


def run_deck(path, patch: str | None = None, out_dir=None, nx=None, ny=None, nz=None, precision: int = 0,
             steps: int | None = None) -> list:
    """Run a channel deck on the GPU: num_iter steps (or ``steps``),
    progress.out rows every print_control steps, rst.bin and
    field_info0.dat at the end.  Returns the rows as text columns."""
    from .io import write_field_info, write_restart
    from .solver import ChannelSolver, progress_row, progress_text

    deck = read_deck(path)
    if patch:
        deck = apply_patch(deck, patch)
    run = case_from_deck(deck, nx=nx, ny=ny, nz=nz, precision=precision)
    c = run.case
    nsteps = run.num_iter if steps is None else steps
    rows = []
    with ChannelSolver(c.params, c.grid) as g:
        g.copy_from_field(c.w0)
        g.init_run()
        done = 0
        while done < nsteps:
            n = min(run.print_control - done % run.print_control, nsteps - done)
            g.advance(n)
            done += n
            if done % run.print_control == 0 or done == nsteps:
                rows.append(progress_text(progress_row(g.compute_residual(), c.params)))
        w = g.copy_to_field()
        sc = g.scalars()
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        with open(os.path.join(out_dir, "progress.out"), "w") as f:
            for r in rows:
                f.write(f"{int(r[0]):10d}" + "".join(f"{float(v):20.10E}" for v in r[1:]) + "\n")
        write_restart(os.path.join(out_dir, "rst.bin"), w, c.grid.ng)
        write_field_info(os.path.join(out_dir, "field_info0.dat"), {"icyc0": sc["icyc"], "time0": sc["time"]})
    return rows


def run_deck_fortran(path, out_dir, patch: str | None = None, precision: int = 0, steps: int | None = None,
                     timeout: float = 600.0) -> str:
    """The same deck through the Fortran driver (streams_mi355x_run, the
    bind(c) layer): the case record is written to out_dir/case.bin and the
    driver writes rst.bin / field_info0.dat there.  Returns its output."""
    from .fortran_case import run_driver, write_case

    deck = read_deck(path)
    if patch:
        deck = apply_patch(deck, patch)
    run = case_from_deck(deck, precision=precision)
    os.makedirs(out_dir, exist_ok=True)
    case = os.path.join(out_dir, "case.bin")
    write_case(case, run.case.params, run.case.grid, run.case.w0, run.num_iter if steps is None else steps)
    return run_driver(case, os.path.join(out_dir, "result.bin"), timeout=timeout, files_dir=out_dir)


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="Run a channel singleideal.ini deck on the GPU")
    ap.add_argument("deck")
    ap.add_argument("--patch", default=None, help="tests.ini patch: 'section,key=value;...'")
    ap.add_argument("--out", default=".", help="directory for progress.out, rst.bin, field_info0.dat")
    ap.add_argument("--steps", type=int, default=None, help="steps instead of the deck's num_iter")
    ap.add_argument("--mixed", action="store_true", help="the mixed FP32-flux / FP64-state form")
    ap.add_argument("--fortran", action="store_true", help="run through the Fortran driver (bind(c) layer)")
    a = ap.parse_args(argv)
    if a.fortran:
        print(run_deck_fortran(a.deck, a.out, a.patch, precision=1 if a.mixed else 0, steps=a.steps), end="")
        return
    rows = run_deck(a.deck, a.patch, a.out, precision=1 if a.mixed else 0, steps=a.steps)
    print(" ".join(rows[-1]))


if __name__ == "__main__":
    main()
