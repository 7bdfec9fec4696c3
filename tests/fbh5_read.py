"""Reads FBH5 files back for the tests with HDF5's own h5dump (h5py is not
installed): the attributes, the data set's shape and chunks, its filters
and its float32 data."""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def available() -> bool:
    return os.path.exists(H5DUMP)


def _attrs(block: str) -> dict:
    out = {}
    for m in re.finditer(r'ATTRIBUTE "([^"]+)" \{(.*?)DATA \{\s*\(0\): (.*?)\s*\}', block, re.S):
        name, body, val = m.group(1), m.group(2), m.group(3).strip()
        if val.startswith('"'):
            strs = re.findall(r'"([^"]*)"', val)
            out[name] = strs if len(strs) > 1 else strs[0]
        elif "H5T_IEEE_F64" in body or "H5T_IEEE_F32" in body:
            out[name] = float(val)
        else:
            out[name] = int(val)
    return out


def read(path: str):
    """(file attributes, data set attributes, info, data as float32 array)."""
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-H", "-A", "-p", path], capture_output=True,
                         text=True, check=True).stdout
    ds_at = txt.index('DATASET "data"')
    file_part, ds_part = txt[:ds_at], txt[ds_at:]
    shape = tuple(int(x) for x in re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", ds_part)
                  .group(1).split(","))
    chunk = tuple(int(x) for x in re.search(r"CHUNKED \( ([^)]*) \)", ds_part).group(1).split(","))
    filters = re.search(r"FILTERS \{\s*(.*?)\s*\}", ds_part, re.S).group(1).split()[0]
    dtype = re.search(r"DATATYPE\s+(\S+)", ds_part).group(1)
    info = {"shape": shape, "chunk": chunk, "filters": filters, "dtype": dtype}
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "data.bin")
        subprocess.run([H5DUMP, "-d", "data", "-b", "LE", "-o", out, path], capture_output=True,
                       check=True)
        data = np.fromfile(out, dtype=np.float32) if os.path.exists(out) else np.zeros(0, np.float32)
    return _attrs(file_part), _attrs(ds_part), info, data.reshape(shape) if data.size else data
