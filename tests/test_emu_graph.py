"""The batched form's whole kernel sequence on the CPU (dnaln_amd/bin/graph_emu).

graph_emu is the library built over a CPU emulation of HIP (csrc/emu/hip): every thread of a
block runs as a fiber, waves step together at their collectives (shuffles, DPP, ballot,
readfirstlane, ds_bpermute) and barriers, device memory is host memory allocated at the exact
size asked for, and the build carries AddressSanitizer, so any access outside a buffer, LDS array
or kernel argument aborts the run. Both launch forms run here: the kernels one by one, and the
captured graph replayed (this build keeps graph capture). Records must equal the reference
binary's records of the golden sets; small read ranges keep each case to seconds. GRAPH selects
the batched form: 0 the host-synchronized form, 2 the device-count sequence launched one by one,
1 that sequence captured as a graph (the device build refuses 1 and 2)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dnaln_amd import abi, api
from tests import golden_io
from tests.golden import sets

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "dnaln_amd" / "bin" / "graph_emu"

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="graph_emu not built (python -m dnaln_amd.build emu)")

_DIRS: dict = {}


def _inputs(name: str, tmp_root: Path) -> Path:
    """The set's genome, reads and parameters as the raw arrays graph_emu reads."""
    if name in _DIRS:
        return _DIRS[name]
    spec = sets.SETS[name]
    g = sets.genome(spec)
    rb = api.prepare(sets.reads(spec, g))
    d = tmp_root / name
    d.mkdir(parents=True, exist_ok=True)
    g.codes.astype(np.uint8).tofile(d / "genome.u8")
    np.stack([g.contig_off.astype(np.uint64), g.contig_len.astype(np.uint64)], 1).tofile(d / "contigs.u64")
    rb.codes.tofile(d / "codes.u8")
    rb.offs.astype(np.uint64).tofile(d / "offs.u64")
    rb.lens.astype(np.int32).tofile(d / "lens.i32")
    rb.fnv.astype(np.uint32).tofile(d / "fnv.u32")
    mp = sets.map_params(spec)
    np.array([mp[k] for k in ("k", "w", "max_occ", "match", "mismatch", "gap_open", "gap_ext", "bw", "max_cand",
                              "min_chain", "min_aln_score")], np.int32).tofile(d / "params.i32")
    _DIRS[name] = d
    return d


@pytest.fixture(scope="module")
def tmp_root(tmp_path_factory):
    return tmp_path_factory.mktemp("graph_emu")


def _run(name: str, tmp_root: Path, ranges: list[tuple[int, int]], graph: int = 1, split=(0, 0), apr: int = 0,
         dir_budget: int = 0, walk: int = 0, poison: int = 0xCD, mode: int = abi.MODE_BATCHED) -> list[str]:
    """poison: the byte every new device allocation starts with. 0xCD makes a stale int32 count
    negative; 0x01 makes it large and positive, as reused device memory can hold."""
    d = _inputs(name, tmp_root)
    cmd = [str(EXE), str(d), str(mode), str(graph), str(split[0]), str(split[1]), str(apr),
           str(dir_budget), str(walk)] + [f"{lo}:{hi}" for lo, hi in ranges]
    env = dict(os.environ, ASAN_OPTIONS="detect_stack_use_after_return=0:detect_leaks=0", HIP_EMU_CUS="8",
               HIP_EMU_POISON=str(poison))
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
    assert p.returncode == 0, p.stderr[-4000:]
    spec = sets.SETS[name]
    g = sets.genome(spec)
    gold = golden_io.records(name)
    for i, (lo, hi) in enumerate(ranges):
        recs = np.fromfile(d / f"rec_{i}.bin", abi.RECORD_DTYPE)
        cig = np.fromfile(d / f"cig_{i}.u32", np.uint32)
        assert len(recs) == hi - lo
        bad = [j for j in range(hi - lo) if api.record_fields(recs[j], cig, g.names) != gold[lo + j][1:]]
        assert bad == [], (name, lo, hi, bad[:5])
    return p.stdout.splitlines()


def test_graph_replays_and_new_plans_give_reference_records(tmp_root):
    """One range run twice (the second run replays the captured graph), then a range of another
    size (a new plan, a second graph)."""
    out = _run("uniform4m_150bp_10k", tmp_root, [(0, 300), (0, 300), (40, 250)])
    assert len(out) == 3


def test_global_path_and_wave_chain_in_the_graph(tmp_root):
    """Repeat-rich reads with many anchors: the global path, its scan and the wave chain."""
    _run("edge_repeats_bw32", tmp_root, [(0, 300)])


def test_direct_launches_give_reference_records(tmp_root):
    _run("edge_uniform_bw64", tmp_root, [(0, 200)], graph=2)


def test_anchor_overflow_runs_again(tmp_root):
    """One anchor per read and no fused buckets: the first run overflows the anchor buffer, the
    host grows it and captures and runs again."""
    out = _run("edge_repeats_bw32", tmp_root, [(0, 150)], split=(-1, -1), apr=1)
    f = out[0].split()
    assert int(f[f.index("anchors") + 1]) > 150


def test_cut_traceback_and_window_walk(tmp_root):
    """A small direction budget cuts the traceback fill into several launches; the walk through
    LDS windows of flag rows."""
    _run("edge_uniform_bw16", tmp_root, [(0, 300)], dir_budget=1 << 20, walk=1)


def test_uneven_lengths_and_narrow_band(tmp_root):
    """Reads of uneven length (length sort, per-wave direction slots) at bw 5."""
    _run("edge_twins_sr_bw5", tmp_root, [(0, 300)])


@pytest.mark.parametrize("poison", [0x01, 0x7F])
def test_anchor_overflow_with_stale_device_memory(tmp_root, poison):
    """The run that overflows the anchor buffer still runs every later stage before the host runs
    it again; those stages must not read a count the overflowed chain stage left unwritten. With
    device memory holding large positive stale values, a read's unwritten candidate count once sent
    the rank-task stage far outside its buffers."""
    out = _run("edge_repeats_bw32", tmp_root, [(0, 150)], graph=2, split=(-1, -1), apr=1, poison=poison)
    f = out[0].split()
    assert int(f[f.index("anchors") + 1]) > 150


@pytest.mark.parametrize("form", [(abi.MODE_BATCHED, 2), (abi.MODE_BATCHED, 0), (abi.MODE_PLAIN, 0)],
                         ids=["batched-device-counts", "batched", "plain"])
@pytest.mark.parametrize("name", ["uniform4m_150bp_10k", "edge_twins_sr_bw5", "edge_uniform_bw64"])
def test_stale_device_memory_every_form(tmp_root, name, form):
    """Every stage of every form on device memory that starts as large positive values: no stage
    may consume a value it did not write first."""
    _run(name, tmp_root, [(0, 200), (30, 230)], graph=form[1], poison=0x01, mode=form[0])
