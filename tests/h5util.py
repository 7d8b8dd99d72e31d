"""h5dump-based readers for the HDF5 tests (independent of libicempm_h5)."""
import re
import shutil
import subprocess

import numpy as np

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def header(path):
    return subprocess.run([H5DUMP, "-H", str(path)], check=True, stdout=subprocess.PIPE,
                          text=True).stdout


def structure(path):
    """(dataset, type, rank) and (attribute, type) entries of a file header."""
    text = header(path)
    out = set()
    for m in re.finditer(r'DATASET "(\w+)" \{\s*DATATYPE\s+(\w+)\s*DATASPACE\s+SIMPLE \{ \(([^)]*)\)', text):
        out.add(("dataset", m.group(1), m.group(2), len(m.group(3).split(","))))
    for m in re.finditer(r'ATTRIBUTE "(\w+)" \{\s*DATATYPE\s+(\w+)', text):
        out.add(("attribute", m.group(1), m.group(2)))
    return out


def attr_values(path, attr):
    """The values of an attribute as h5dump prints them."""
    text = subprocess.run([H5DUMP, "-a", attr, str(path)], check=True, stdout=subprocess.PIPE,
                          text=True).stdout
    body = text[text.index("DATA {"):]
    return [float(v) for v in re.findall(r"\(\d+\):\s*([-+0-9.eE]+)", body)]


def binary(path, dataset, dtype, tmp_path):
    out = tmp_path / (dataset.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", dataset, "-b", "LE", "-o", str(out), str(path)], check=True,
                   stdout=subprocess.DEVNULL)
    return np.fromfile(out, dtype=dtype)
