"""Digitize the reference's attenuation-accuracy figure into a JSON fixture.

Input: ``figs/attenuation_modeling.svg`` of the reference (a matplotlib SVG, read as text).
Output: ``tests/golden/attenuation_modeling_reference.json`` with, for each of the 15 panels
(alpha0 in {0.25, 0.5, 0.75} dB/(MHz^gamma cm) by rows, gamma in {0.4, 0.7, 1.0, 1.3, 1.6} by
columns, as the figure's row and column labels give them), the 8 "FW simulation" markers as
(frequency in MHz, attenuation in Np/m) and their relative error against alpha0 f^gamma.

Each panel's pixel -> data map is fitted from its tick marks and tick labels. The panel's
"Analytical" curve is checked against alpha0 f^gamma to bound the digitization error.

Usage: python tests/golden/digitize_attenuation_svg.py [path/to/attenuation_modeling.svg]
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np

NP_PER_DB_CM = 100.0 / (20.0 * np.log10(np.e))  # dB/cm -> Np/m
ALPHAS = [0.25, 0.5, 0.75]
GAMMAS = [0.4, 0.7, 1.0, 1.3, 1.6]


def _ticks(block: str, axis: str):
    """(pixel, value) pairs of the ticks of one axis inside an axes block."""
    out = []
    for m in re.finditer(rf'<g id="{axis}tick_\d+">(.*?)</g>\s*</g>\s*</g>', block, re.S):
        body = m.group(1)
        use = re.search(r'<use xlink:href="#m[0-9a-f]+" x="([\d.]+)" y="([\d.]+)"', body)
        lab = re.search(r"<!-- ([^>]*?) -->", block[m.end():m.end() + 400]) or re.search(r"<!-- ([^>]*?) -->", body)
        if use and lab:
            try:
                val = float(lab.group(1).replace("−", "-"))
            except ValueError:
                continue
            out.append((float(use.group(1 if axis == "x" else 2)), val))
    return out


def _fit(pairs):
    px = np.array([p for p, _ in pairs])
    val = np.array([v for _, v in pairs])
    a, b = np.polyfit(px, val, 1)
    return lambda q: a * np.asarray(q) + b


def digitize(svg_text: str) -> dict:
    starts = [m.start() for m in re.finditer(r'<g id="axes_\d+">', svg_text)]
    starts.append(len(svg_text))
    panels = []
    for n in range(15):
        block = svg_text[starts[n] : starts[n + 1]]
        xt = _ticks(block, "x")
        yt = _ticks(block, "y")
        fx, fy = _fit(xt), _fit(yt)
        legend = block.find('<g id="legend_')
        data = block[:legend] if legend > 0 else block
        markers = re.findall(r'<use xlink:href="#m[0-9a-f]+" x="([\d.]+)" y="([\d.]+)"', data)
        # tick marks are <use> too: keep the ones that are not tick positions
        tick_px = {round(p, 3) for p, _ in xt} | {round(p, 3) for p, _ in yt}
        pts = [(float(x), float(y)) for x, y in markers
               if not (round(float(x), 3) in tick_px and round(float(y), 3) in {232.0, 472.0, 712.0} | tick_px)]
        pts = [p for p in pts if p not in [(float(x), float(y)) for x, y in markers[:0]]]
        # the 8 simulation markers are the last 8 <use> before the legend
        sim = [(float(x), float(y)) for x, y in markers[-8:]]
        f = fx([p[0] for p in sim])
        att = fy([p[1] for p in sim])
        alpha0, gamma = ALPHAS[n // 5], GAMMAS[n % 5]
        target = alpha0 * f**gamma * NP_PER_DB_CM
        # analytical curve: the path with many vertices and no markers
        curves = re.findall(r'<path d="(M [^"]+)"[^>]*clip-path', data)
        dig_err = None
        for d in curves:
            xy = np.array([[float(v) for v in seg.split()] for seg in re.findall(r"[ML] ([\d.]+ [\d.]+)", d)])
            if xy.shape[0] > 20:
                cf, ca = fx(xy[:, 0]), fy(xy[:, 1])
                ok = cf > 0.5
                ref = alpha0 * cf[ok] ** gamma * NP_PER_DB_CM
                dig_err = float(np.max(np.abs(ca[ok] - ref)) / np.max(ref))
                break
        panels.append({
            "panel": n + 1,
            "alpha0_db_mhz_cm": alpha0,
            "gamma": gamma,
            "freq_mhz": [round(float(v), 4) for v in f],
            "sim_np_per_m": [round(float(v), 4) for v in att],
            "rel_err": [round(float(v), 5) for v in att / target - 1],
            "digitization_error_of_analytical_curve": dig_err,
        })
    return {
        "source": "figs/attenuation_modeling.svg of pinton-lab/fullwave25 v1.0.16",
        "units": {"alpha0": "dB/(MHz^gamma cm)", "attenuation": "Np/m", "frequency": "MHz"},
        "panels": panels,
    }


def main() -> None:
    svg = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference/figs/attenuation_modeling.svg")
    out = Path(__file__).resolve().parent / "attenuation_modeling_reference.json"
    result = digitize(svg.read_text())
    out.write_text(json.dumps(result, indent=1) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
