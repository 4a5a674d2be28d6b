"""Helpers to run the drop-in driver (tlbm_amd/tlbm_exec) in a directory."""
import os
import subprocess

import numpy as np

from tlbm_amd import io_formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXEC = os.path.join(ROOT, "tlbm_amd", "tlbm_exec")
RESTART = os.path.join(ROOT, "tlbm_amd", "tlbm_restart")


def require(path):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build with `make -C tlbm_amd all`")
    return path


def run_exec(directory, storage=None, env=None, timeout=300):
    e = dict(os.environ)
    if storage:
        e["TLBM_STORAGE"] = storage
    if env:
        e.update(env)
    return subprocess.run([require(EXEC)], cwd=directory, capture_output=True, text=True,
                          env=e, timeout=timeout)


def case_dir(tmp_path, params, nd, parts=True):
    d = str(tmp_path)
    io_formats.write_case(d, params, nd, None if parts else None)
    if not parts:
        os.remove(os.path.join(d, "parts.lbm"))
    return d


def read_dump(directory, k, dtype):
    return {f: io_formats.read_bdat(directory, f, k, dtype) for f in io_formats.DUMP_FIELDS}


def read_raw(directory, name, dtype):
    return np.fromfile(os.path.join(directory, name), dtype=dtype)
