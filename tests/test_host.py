"""Host-side behaviour that needs no GPU: API validation, chunk planning,
workload generators, and the C ABI of the HIP library (loads and exports
every declared symbol; no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from paragram_amd import _api, _chunking, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- validation
# Messages and exception types of paragram/_api.py:16-62, checked on CPU
# tensors (the device check is the first failure there).

def test_points_type_and_device():
    with pytest.raises(TypeError, match="points must be a torch.Tensor"):
        _api._validate_points(np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError, match="points must be on a CUDA device"):
        _api._validate_points(torch.zeros(3, 3))


class FakeCuda(torch.Tensor):
    """A CPU tensor that reports is_cuda, to reach the later checks."""

    @property
    def is_cuda(self):
        return True


def fake(t):
    return t.as_subclass(FakeCuda)


def test_points_dtype_and_shape():
    with pytest.raises(ValueError, match="points must have dtype torch.float32"):
        _api._validate_points(fake(torch.zeros(3, 3, dtype=torch.float64)))
    with pytest.raises(ValueError, match=re.escape("points must have shape (N, 3)")):
        _api._validate_points(fake(torch.zeros(3, 2)))
    with pytest.raises(ValueError, match=re.escape("points must have shape (N, 3)")):
        _api._validate_points(fake(torch.zeros(3)))
    out = _api._validate_points(fake(torch.zeros(4, 3)))
    assert out.is_contiguous()


def test_weights_checks():
    dev = torch.device("cpu")
    with pytest.raises(TypeError, match="weights must be a torch.Tensor"):
        _api._validate_weights([1.0], 1, dev)
    with pytest.raises(ValueError, match="weights must be on a CUDA device"):
        _api._validate_weights(torch.zeros(3), 3, dev)
    with pytest.raises(ValueError, match="weights must have dtype torch.float32"):
        _api._validate_weights(fake(torch.zeros(3, dtype=torch.float64)), 3, dev)
    with pytest.raises(ValueError, match=re.escape("weights must have shape (N,) or (N, 1)")):
        _api._validate_weights(fake(torch.zeros(4)), 3, dev)
    with pytest.raises(ValueError, match=re.escape("weights must have shape (N,) or (N, 1)")):
        _api._validate_weights(fake(torch.zeros(3, 2)), 3, dev)
    assert tuple(_api._validate_weights(fake(torch.zeros(3)), 3, dev).shape) == (3, 1)
    assert tuple(_api._validate_weights(fake(torch.zeros(3, 1)), 3, dev).shape) == (3, 1)


def test_initial_guess_checks():
    dev = torch.device("cpu")
    assert _api._validate_initial_guesses(None, None, 3, dev) == (None, None)
    with pytest.raises(ValueError, match="must be provided together"):
        _api._validate_initial_guesses(fake(torch.zeros(3, dtype=torch.int32)), None, 3, dev)
    with pytest.raises(ValueError, match="initial guess tensors must be on a CUDA device"):
        _api._validate_initial_guesses(torch.zeros(3, dtype=torch.int32), torch.zeros(4, dtype=torch.int32), 3, dev)
    with pytest.raises(ValueError, match="must have dtype torch.int32"):
        _api._validate_initial_guesses(fake(torch.zeros(3)), fake(torch.zeros(4, dtype=torch.int32)), 3, dev)
    with pytest.raises(ValueError, match=re.escape("initial_guesses_offsets must have shape (N + 1,)")):
        _api._validate_initial_guesses(fake(torch.zeros(3, dtype=torch.int32)),
                                       fake(torch.zeros(3, dtype=torch.int32)), 3, dev)


def test_public_api_signatures():
    import inspect

    import paragram_amd

    sig = inspect.signature(paragram_amd.voronoi_diagram)
    assert list(sig.parameters) == ["points", "knn", "debug", "initial_guesses", "initial_guesses_offsets",
                                    "leaf_size"]
    assert sig.parameters["knn"].default == 8 and sig.parameters["leaf_size"].default == -1
    sig = inspect.signature(paragram_amd.power_diagram)
    assert list(sig.parameters)[:2] == ["points", "weights"]
    d = paragram_amd.Diagram(adjacency=1, offsets=2, status=3)
    with pytest.raises(Exception):
        d.adjacency = 5  # frozen dataclass


def test_api_rejects_cpu_points_before_loading_engine():
    import paragram_amd

    with pytest.raises(ValueError, match="points must be on a CUDA device"):
        paragram_amd.voronoi_diagram(torch.zeros(5, 3))


# ---------------------------------------------------------------- chunking
# Table of (total_blocks, full_bytes, free, env) -> plan, following
# chunking.h:16-56 line by line.

@pytest.mark.parametrize("total,full,free,env,expect", [
    (100, 1000, 10_000, {}, (100, 1)),                                  # fits: one chunk
    (100, 1000, 1000, {}, (50, 2)),                                     # budget 500 -> 2 chunks
    (100, 1000, 1000, {"PARAGRAM_MEM_FRAC": "0.25"}, (25, 4)),          # budget 250 -> 4 chunks
    (100, 1000, 1000, {"PARAGRAM_MEM_FRAC": "0.99"}, (50, 2)),          # out of (0.05, 0.95): ignored
    (100, 1000, 1000, {"PARAGRAM_MEM_FRAC": "abc"}, (50, 2)),           # unparsable -> 0 -> ignored
    (100, 1000, 10_000, {"PARAGRAM_NUM_CHUNKS": "4"}, (25, 4)),         # forced
    (15625, 0, 0, {"PARAGRAM_NUM_CHUNKS": "4"}, (3907, 4)),             # C2: 1M cells in 64-cell blocks
    (10, 1000, 10_000, {"PARAGRAM_NUM_CHUNKS": "7"}, (2, 5)),           # fewer chunks than forced
    (3, 1000, 10_000, {"PARAGRAM_NUM_CHUNKS": "50"}, (1, 3)),           # clamp to >= 1 block
    (100, 1000, 10_000, {"PARAGRAM_NUM_CHUNKS": "0"}, (100, 1)),        # 0 is not forcing
    (100, 1000, 10_000, {"PARAGRAM_NUM_CHUNKS": "-3"}, (100, 1)),
    (100, 1000, 0, {}, (100, 1)),                                       # no free memory info: budget 0
])
def test_chunk_plan_table(total, full, free, env, expect):
    p = _chunking.plan_chunks(total, 64, full, free, env)
    assert (p.chunk_blocks, p.num_chunks) == expect
    assert p.chunk_threads == p.chunk_blocks * 64


# ---------------------------------------------------------------- workloads

def test_workload_shapes_and_determinism():
    a = workloads.uniform(1000, 0)
    assert a.dtype == np.float32 and a.shape == (1000, 3)
    assert np.array_equal(a, workloads.uniform(1000, 0))
    p, w = workloads.clustered_power(10_000, 2)
    assert p.shape == (10_000, 3) and w.shape == (10_000,)
    assert p.min() >= 0 and p.max() <= 1 and w.min() >= 0 and w.max() <= 3e-5
    sizes = workloads.lidar_sizes(3)
    assert sizes.shape == (256,) and sizes.min() >= 20_000 and sizes.max() <= 200_000
    clouds = workloads.lidar_batch(3, max_clouds=2)
    assert [c.shape[0] for c in clouds] == list(sizes[:2])
    s = workloads.oracle_sample(1_000_000, 4, 1000)
    assert s.shape == (1000,) and np.all(np.diff(s) > 0)


# ---------------------------------------------------------------- C ABI

def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "paragram_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(pg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from paragram_amd import _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    lib = ctypes.CDLL(_engine.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert lib.pg_abi_version() == 2


def test_library_argument_errors_without_gpu():
    """Calls that fail argument checks before touching the device."""
    from paragram_amd import _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    lib = _engine.HipLibrary()
    out = ctypes.c_uint64(0)
    assert lib.lib.pg_index_workspace_bytes(-1, ctypes.byref(out)) == 1
    assert b"non-negative" in lib.lib.pg_last_error()
    assert lib.lib.pg_index_workspace_bytes(1 << 27, ctypes.byref(out)) == 1
    assert lib.lib.pg_index_workspace_bytes(1000, ctypes.byref(out)) == 0 and out.value > 1000 * 16
    assert lib.lib.pg_compute_cells(10, None, 0, 20, None, None, 0, 16, None, 0, None, None, None, None, None, 0,
                                    1, None, None) == 1
    assert b"out of bounds" in lib.lib.pg_last_error()
    # the row pool and the cell scratch are sized per cell (16 bytes of
    # scratch: the FP64 flag and three queues, so no queue can fill; 48 plane
    # owners for one cell in 32, and 64 more, queued for tier m; a 20-byte
    # record for one removal test per two cells, and 1024 more, that tier 0
    # took by its FP32 sign)
    assert lib.lib.pg_row_pool_elems(1000, ctypes.byref(out)) == 0 and out.value == 32 * 1000 + 65536
    assert lib.lib.pg_cell_scratch_bytes(1000, ctypes.byref(out)) == 0 and out.value == 256 + 16 * 1000 + 4 * 48 * (1000 // 32 + 64) + 20 * (1000 // 2 + 1024)
    # the index layout (host only): increasing offsets inside the workspace
    offs = (ctypes.c_uint64 * 6)()
    assert lib.lib.pg_index_layout(1000, offs) == 0
    assert list(offs) == sorted(offs) and offs[0] == 0
    assert lib.lib.pg_index_workspace_bytes(1000, ctypes.byref(out)) == 0 and out.value == offs[5]


# ---------------------------------------------------------------- batches

def test_batch_quads_host_helper():
    from paragram_amd import _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    lib = _engine.HipLibrary()
    off = np.array([0, 0, 5, 6, 14, 14], np.int64)  # sizes 0, 5, 1, 8, 0
    q = np.zeros(6, np.int64)
    p = ctypes.c_void_p
    assert lib.lib.pg_batch_quads(5, p(off.ctypes.data), 0, 14, p(q.ctypes.data)) == 0
    assert list(q) == [0, 0, 2, 3, 5, 5]
    # a sub-range cutting clouds: cells [3, 9) -> cloud 1 has 2, cloud 2 has 1, cloud 3 has 3
    assert lib.lib.pg_batch_quads(5, p(off.ctypes.data), 3, 9, p(q.ctypes.data)) == 0
    assert list(q) == [0, 0, 1, 2, 3, 3]
    assert lib.lib.pg_batch_quads(5, p(off.ctypes.data), 3, 15, p(q.ctypes.data)) == 1
    assert b"out of bounds" in lib.lib.pg_last_error()


def test_batch_offsets_rejected_before_any_device_work():
    from paragram_amd import _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    lib = _engine.HipLibrary()
    p = ctypes.c_void_p
    bad = np.array([0, 5, 3, 10], np.int64)
    assert lib.lib.pg_build_batch_index(10, 3, None, None, p(bad.ctypes.data), -1, None, 0, None) == 1
    assert b"non-decreasing" in lib.lib.pg_last_error()
    bad = np.array([0, 5, 9], np.int64)
    assert lib.lib.pg_build_batch_index(10, 2, None, None, p(bad.ctypes.data), -1, None, 0, None) == 1
    assert b"end at n" in lib.lib.pg_last_error()
    out = ctypes.c_uint64(0)
    assert lib.lib.pg_batch_workspace_bytes(10, 5000, ctypes.byref(out)) == 1


def test_batch_api_validation():
    with pytest.raises(ValueError, match="start at 0, end at N"):
        _api._validate_cloud_offsets([0, 3, 2, 7], 7)
    with pytest.raises(ValueError, match="start at 0, end at N"):
        _api._validate_cloud_offsets([1, 7], 7)
    with pytest.raises(ValueError, match="1-D integer"):
        _api._validate_cloud_offsets([0.0, 7.0], 7)
    assert list(_api._validate_cloud_offsets(torch.tensor([0, 2, 7]), 7)) == [0, 2, 7]


def test_lazy_build_on_first_import(monkeypatch, tmp_path, capsys):
    """A missing library is built on first load (as the reference's JIT build
    on first import, paragram/_backend.py:147-190), a present one is loaded
    as is, and PARAGRAM_NO_BUILD=1 leaves the engine unavailable instead.
    The builder is replaced by a copy of the built library (the real hipcc
    build is what __graft_entry__.build() runs)."""
    import shutil

    from paragram_amd import _backend, _build, _engine

    if not os.path.exists(_engine.LIB_PATH):
        pytest.skip("libparagram_hip.so not built")
    monkeypatch.setattr(_backend, "_load_error", _backend._load_error)
    calls = []

    def fake_build(verbose=None, lib_path=_build.LIB_PATH):
        calls.append(lib_path)
        shutil.copy(_engine.LIB_PATH, lib_path)
        return lib_path

    monkeypatch.setattr(_build, "build", fake_build)
    monkeypatch.delenv("PARAGRAM_NO_BUILD", raising=False)
    target = str(tmp_path / "libparagram_hip.so")
    be = _backend._load(target)
    assert be is not None and calls == [target]
    out = capsys.readouterr().out
    assert "building the HIP library for gfx950" in out and "HIP library ready in" in out
    calls.clear()
    assert _backend._load(target) is not None and calls == []
    monkeypatch.setenv("PARAGRAM_NO_BUILD", "1")
    assert _backend._load(str(tmp_path / "missing.so")) is None
    assert "PARAGRAM_NO_BUILD" in _backend._load_error and calls == []
