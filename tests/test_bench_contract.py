"""bench.py's command-line contract and its byte accounting, on the CPU (no GPU is touched)."""
import importlib.util
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message(capsys):
    b = _bench()
    with pytest.raises(SystemExit) as e:
        b.main(["--gpus", "2"])
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err


def test_algorithmic_bytes_of_the_100_box_products():
    """y = A p on the 100^3 box: 8 B per value, one int32 node index per 3x3 block, 12 B per
    64-node slice plus the end offset, x read and y written once (16 B per row), and the fused
    r0.y read (8 B per row); the CRS product: 12 B per nonzero, int64 row pointers, x and y."""
    b = _bench()
    n, nnz = 3_090_903, 245_438_109
    assert nnz % 9 == 0 and n % 3 == 0
    k1 = b.spmv_bytes(2, n, nnz) + 8.0 * n
    assert k1 == 2_146_963_344
    assert b.spmv_bytes(1, n, nnz) == b.spmv_bytes(4, n, nnz) == 12.0 * nnz + 8.0 * (n + 1) + 16.0 * n
    # FP32 values: half the value bytes, FP32 vectors
    assert b.spmv_bytes(3, n, nnz) == 4.0 * nnz + 4.0 * (nnz // 9) + 12.0 * ((n // 3 + 63) // 64) + 8 + 8.0 * n
