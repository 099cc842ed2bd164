"""Persistent proxy-point cache (tsne_proxy_cache_info, host code, no device needed).

Each check runs in a fresh process (the table is also cached per process): the first
process generates the table and writes the cache file, a second one reads it back and must
get the same table (digest); a corrupted or mismatching file is ignored and rewritten; the
reference's switch TSNE_H2PACK_NO_PROXY_CACHE=1 turns the cache off.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tsne-h2pack-python_amd", "tsne_h2pack", "_lib", "libtsne_h2pack.so")

PROBE = r"""
import ctypes, json, sys
lib = ctypes.CDLL(sys.argv[1])
lib.tsne_proxy_cache_info.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_int)]
buf = ctypes.create_string_buffer(1024)
dg, fd = ctypes.c_ulonglong(), ctypes.c_int()
rc = lib.tsne_proxy_cache_info(int(sys.argv[2]), float(sys.argv[3]), buf, 1024, ctypes.byref(dg), ctypes.byref(fd))
print(json.dumps({"rc": rc, "path": buf.value.decode(), "digest": dg.value, "from_disk": fd.value}))
"""


def _probe(env, dim=2, tol=1e-2):
    r = subprocess.run([sys.executable, "-c", PROBE, LIB, str(dim), str(tol)], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.fixture
def env(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    e = dict(os.environ)
    e["TSNE_H2PACK_CACHE_DIR"] = str(tmp_path)
    e.pop("TSNE_H2PACK_NO_PROXY_CACHE", None)
    return e


def test_generated_then_read_back(env, tmp_path):
    a = _probe(env)
    assert a["rc"] == 0 and a["from_disk"] == 0
    assert a["path"].startswith(str(tmp_path)) and os.path.exists(a["path"])
    with open(a["path"], "rb") as f:
        head = f.read(12)
    assert head[:8] == b"TSNEPRXY" and int.from_bytes(head[8:12], "little") == 1
    b = _probe(env)
    assert b["from_disk"] == 1 and b["digest"] == a["digest"] and b["path"] == a["path"]


def test_corrupted_file_is_replaced(env):
    a = _probe(env)
    with open(a["path"], "r+b") as f:
        f.seek(200)
        byte = f.read(1)
        f.seek(200)
        f.write(bytes([byte[0] ^ 0xFF]))
    b = _probe(env)
    assert b["from_disk"] == 0 and b["digest"] == a["digest"]  # regenerated, same table
    c = _probe(env)
    assert c["from_disk"] == 1 and c["digest"] == a["digest"]  # and rewritten intact


def test_truncated_or_foreign_file_is_replaced(env):
    a = _probe(env)
    with open(a["path"], "r+b") as f:
        f.truncate(100)
    assert _probe(env)["from_disk"] == 0
    with open(a["path"], "wb") as f:
        f.write(b"not a proxy table")
    assert _probe(env)["from_disk"] == 0
    assert _probe(env)["from_disk"] == 1


def test_tables_are_keyed_by_dimension_and_tolerance(env):
    a2 = _probe(env, 2, 1e-2)
    a3 = _probe(env, 3, 3e-2)
    b2 = _probe(env, 2, 1e-3)
    assert len({a2["path"], a3["path"], b2["path"]}) == 3
    assert len({a2["digest"], a3["digest"], b2["digest"]}) == 3
    assert _probe(env, 3, 3e-2)["from_disk"] == 1


def test_cache_can_be_turned_off(env, tmp_path):
    env["TSNE_H2PACK_NO_PROXY_CACHE"] = "1"
    a = _probe(env)
    assert a["path"] == "" and a["from_disk"] == 0
    assert not any(p.name.endswith(".bin") for p in tmp_path.iterdir())


def test_invalid_arguments(env):
    assert _probe(env, 4)["rc"] == -1
