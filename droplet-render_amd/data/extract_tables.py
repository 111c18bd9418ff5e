"""Regenerate the numeric tables the renderer carries as data.

Two tables come from droplet-render and are used unchanged (they are data, not
code, and are redistributed under their BSD-3 licences, reproduced in the
header of every generated file):

* the tabulated Mie phase function of cloud droplets, 256 bins of RGB pdf and
  256 bins of RGB cdf, indexed by acos(-cos theta)/pi
  (reference: src/KernelSampler.cpp:8-73 `miepdf`, :76-141 `miecdf`);
* the Hosek-Wilkie 2012 RGB sky dataset: for each of R, G, B, 2 albedos x 10
  turbidities x 6 elevation control points x 9 coefficients, plus 2 x 10 x 6
  radiance control points (reference: src/ArHosekSkyModelData_RGB.h:105-3861,
  `datasetRGB1..3`, `datasetRGBRad1..3`).

The output files hold one C literal per entry, comma separated, so that the
C++/HIP sources include them inside an array initialiser and the Python side
parses them with `load_inc`. The literal text is kept exactly as written in the
source so the float/double values are bit-identical.

Usage: python extract_tables.py /path/to/droplet-render   (reads src/ only)
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

DROPLET_NOTICE = """droplet-render data table. Copyright (c) 2016, jparkkila. All rights reserved.
Redistribution and use in source and binary forms, with or without modification,
are permitted provided that the following conditions are met: 1. Redistributions
of source code must retain the above copyright notice, this list of conditions
and the following disclaimer. 2. Redistributions in binary form must reproduce
the above copyright notice, this list of conditions and the following disclaimer
in the documentation and/or other materials provided with the distribution.
3. Neither the name of the copyright holder nor the names of its contributors
may be used to endorse or promote products derived from this software without
specific prior written permission. THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT
HOLDERS AND CONTRIBUTORS "AS IS" AND ANY EXPRESS OR IMPLIED WARRANTIES ARE
DISCLAIMED (full BSD-3 text: droplet-render LICENSE)."""

HOSEK_NOTICE = """Hosek-Wilkie sky model RGB dataset (Aug 23 dataset).
Copyright (c) 2012 - 2013, Lukas Hosek and Alexander Wilkie. All rights reserved.
Redistribution and use in source and binary forms, with or without modification,
are permitted provided that the following conditions are met: Redistributions of
source code must retain the above copyright notice, this list of conditions and
the following disclaimer. Redistributions in binary form must reproduce the above
copyright notice, this list of conditions and the following disclaimer in the
documentation and/or other materials provided with the distribution. None of the
names of the contributors may be used to endorse or promote products derived from
this software without specific prior written permission. THIS SOFTWARE IS
PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS" AND ANY EXPRESS OR
IMPLIED WARRANTIES ARE DISCLAIMED (full BSD-3 text in the Hosek-Wilkie sources)."""

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?"


def _block(text, name):
    m = re.search(re.escape(name) + r"\s*\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise RuntimeError("table %s not found" % name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return re.findall(NUM, body)


def _write(path, notice, layout, values, per_line):
    with open(path, "w") as f:
        f.write("/*\n" + notice + "\n\nLayout: " + layout + "\n"
                "Generated by extract_tables.py; do not edit.\n*/\n")
        for i in range(0, len(values), per_line):
            f.write(", ".join(values[i:i + per_line]) + ",\n")


def main(ref_root):
    src = os.path.join(ref_root, "src")
    with open(os.path.join(src, "KernelSampler.cpp")) as f:
        ks = f.read()
    pdf = _block(ks, "miepdf")
    cdf = _block(ks, "miecdf")
    assert len(pdf) == 256 * 3 and len(cdf) == 256 * 3, (len(pdf), len(cdf))
    _write(os.path.join(HERE, "mie_pdf_rgb.inc"), DROPLET_NOTICE,
           "float[256][3], bin i covers acos(-cos theta)/pi in [i/256,(i+1)/256) (bin 0 = backscatter)",
           pdf, 3)
    _write(os.path.join(HERE, "mie_cdf_rgb.inc"), DROPLET_NOTICE,
           "float[256][3], running integral of mie_pdf_rgb over the same bins, last entry 1",
           cdf, 3)

    with open(os.path.join(src, "ArHosekSkyModelData_RGB.h")) as f:
        hk = f.read()
    coeffs, rads = [], []
    for ch in (1, 2, 3):
        c = _block(hk, "datasetRGB%d" % ch)
        r = _block(hk, "datasetRGBRad%d" % ch)
        assert len(c) == 2 * 10 * 6 * 9 and len(r) == 2 * 10 * 6, (len(c), len(r))
        coeffs += c
        rads += r
    _write(os.path.join(HERE, "hosek_rgb_coeffs.inc"), HOSEK_NOTICE,
           "double[3 channel][2 albedo][10 turbidity][6 elevation control point][9 coefficient]",
           coeffs, 9)
    _write(os.path.join(HERE, "hosek_rgb_radiance.inc"), HOSEK_NOTICE,
           "double[3 channel][2 albedo][10 turbidity][6 elevation control point]",
           rads, 6)
    print("wrote tables to", HERE)


def load_inc(name):
    """Parse one generated table back into a list of Python floats."""
    with open(os.path.join(HERE, name)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return [float(v.rstrip("f")) for v in re.findall(NUM, text)]


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
