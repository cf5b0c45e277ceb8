"""HTML view of a run report (the JSON written by xplus_device, whose keys
are those of the reference's reports, archived/host/report_io.hpp:165-273).

Plays the part of the reference's scripts/render_report.py: the problem,
the accuracy verdict, host and kernel timing bars, the convergence curve
(log10 of the per-iteration residual, inline SVG) and the iteration table.
One self-contained static page, no scripts.

    python -m xplus_amd.report [static|interactive] report.json report.html
"""
from __future__ import annotations

import html
import json
import math
import sys
from pathlib import Path

HOST_KEYS = ["host_setup", "buffer_h2d", "kernel_total", "buffer_d2h", "verify"]
KERNEL_KEYS = ["pcg_control", "spmv", "init", "dot", "update_xrz", "update_p"]


def _sci(v) -> str:
    return f"{float(v):.3e}"


def _bars(items, total) -> str:
    total = max(total, 1e-12)
    rows = []
    for label, value in items:
        width = max(1.0, 100.0 * value / total) if value > 0 else 0.0
        rows.append(f"<tr><td>{html.escape(label)}</td><td class='axis'><div class='bar' "
                    f"style='width:{width:.2f}%'></div></td><td class='num'>{value:.3f} ms</td></tr>")
    return "<table class='bars'>" + "".join(rows) + "</table>"


def _curve(traces, width=720, height=260) -> str:
    pts = [(t["iteration"], t["residual"]) for t in traces if t.get("residual", 0) > 0]
    if len(pts) < 2:
        return "<p class='note'>No per-iteration trace in this report.</p>"
    step = max(1, len(pts) // 2000)
    pts = pts[::step] + ([pts[-1]] if (len(pts) - 1) % step else [])
    xs = [p[0] for p in pts]
    ys = [math.log10(p[1]) for p in pts]
    x0, x1 = min(xs), max(xs)
    y0, y1 = min(ys), max(ys)
    if y1 - y0 < 1e-12:
        y1 = y0 + 1.0
    pad = 40

    def px(x):
        return pad + (width - 2 * pad) * (x - x0) / max(x1 - x0, 1)

    def py(y):
        return height - pad - (height - 2 * pad) * (y - y0) / (y1 - y0)

    poly = " ".join(f"{px(x):.1f},{py(y):.1f}" for x, y in zip(xs, ys))
    ticks = []
    for e in range(math.floor(y0), math.ceil(y1) + 1):
        if y0 - 1e-9 <= e <= y1 + 1e-9:
            ticks.append(f"<line x1='{pad}' x2='{width - pad}' y1='{py(e):.1f}' y2='{py(e):.1f}' class='grid'/>"
                         f"<text x='4' y='{py(e) + 4:.1f}'>1e{e}</text>")
    return (f"<svg width='{width}' height='{height}' class='curve'>"
            + "".join(ticks)
            + f"<polyline points='{poly}' class='line'/>"
            f"<text x='{pad}' y='{height - 8}'>iteration {x0}</text>"
            f"<text x='{width - pad - 90}' y='{height - 8}'>iteration {x1}</text></svg>")


def _trace_table(traces, head=50, tail=50) -> str:
    if not traces:
        return ""
    rows = traces if len(traces) <= head + tail else traces[:head] + [None] + traces[-tail:]
    body = []
    for t in rows:
        if t is None:
            body.append(f"<tr><td colspan='6' class='note'>... {len(traces) - head - tail} iterations ...</td></tr>")
            continue
        body.append("<tr>" + "".join(f"<td>{v}</td>" for v in
                                     (t["iteration"], _sci(t["alpha"]), _sci(t["beta"]), _sci(t["rz"]),
                                      _sci(t["rr"]), _sci(t["residual"]))) + "</tr>")
    return ("<table class='trace'><tr><th>iteration</th><th>alpha</th><th>beta</th><th>rz</th><th>rr</th>"
            "<th>residual</th></tr>" + "".join(body) + "</table>")


def render_html(data: dict) -> str:
    ds, res = data["dataset"], data["result"]
    host = data.get("host_timing_ms", {})
    kt = data.get("kernel_timing_ms", {})
    traces = data.get("iterations_trace", [])
    meta = data.get("trace_metadata", {})
    verdict = "PASS" if res.get("pass") else "FAIL"
    rows = [("status", res.get("status")), ("iterations", res.get("iterations")),
            ("final r<sup>T</sup>r", _sci(res.get("final_rr", 0))),
            ("final residual", _sci(res.get("final_residual_norm", 0))),
            ("relative residual", _sci(res.get("residual_rel", 0))),
            ("max |x - x_golden|", _sci(res.get("solution_max_abs_diff", 0))),
            ("golden residual", _sci(res.get("golden_residual_norm", 0)))]
    problem = [("dataset", html.escape(str(ds.get("path")))), ("device", html.escape(str(ds.get("xclbin")))),
               ("n / nnz", f"{ds.get('n')} / {ds.get('nnz')}"), ("max_iters", ds.get("max_iters")),
               ("tau", _sci(ds.get("tau", 0)))]

    def kv(items):
        return "<table class='kv'>" + "".join(f"<tr><th>{k}</th><td>{v}</td></tr>" for k, v in items) + "</table>"

    host_items = [(k, float(host.get(k, 0.0))) for k in HOST_KEYS]
    kernel_items = [(k, float(kt.get(k + "_total", 0.0))) for k in KERNEL_KEYS if float(kt.get(k + "_total", 0.0)) > 0]
    note = meta.get("note", "")
    return f"""<!doctype html>
<html><head><meta charset='utf-8'><title>PCG run report</title>
<style>
body {{ font-family: sans-serif; margin: 24px; color: #222; }}
h1 {{ font-size: 20px; }} h2 {{ font-size: 16px; margin-top: 28px; }}
.badge {{ padding: 2px 10px; border-radius: 4px; color: #fff; }}
.PASS {{ background: #2a7d2a; }} .FAIL {{ background: #b22; }}
table {{ border-collapse: collapse; }} td, th {{ padding: 3px 10px; text-align: left; }}
.kv th {{ color: #555; font-weight: normal; }}
.bars .axis {{ width: 420px; background: #eee; }} .bar {{ height: 12px; background: #4a78b8; }}
.num {{ text-align: right; font-variant-numeric: tabular-nums; }}
.trace td {{ font-family: monospace; border-top: 1px solid #eee; }}
.curve .line {{ fill: none; stroke: #4a78b8; stroke-width: 1.5; }} .curve .grid {{ stroke: #ddd; }}
.curve text {{ font-size: 11px; fill: #555; }} .note {{ color: #777; }}
</style></head><body>
<h1>PCG run report <span class='badge {verdict}'>{verdict}</span></h1>
<h2>Problem</h2>{kv(problem)}
<h2>Accuracy</h2>{kv(rows)}
<h2>Host timing</h2>{_bars(host_items, float(host.get("total", 0.0)))}
<p class='note'>total {float(host.get("total", 0.0)):.3f} ms</p>
<h2>Kernel timing</h2>{_bars(kernel_items, sum(v for _, v in kernel_items)) if kernel_items else "<p class='note'>No kernel timing.</p>"}
<p class='note'>timing model: {html.escape(str(kt.get("timing_model", "")))}</p>
<h2>Convergence</h2>{_curve(traces)}
<p class='note'>trace source: {html.escape(str(meta.get("source", "none")))}. {html.escape(note)}</p>
<h2>Iterations</h2>{_trace_table(traces)}
</body></html>
"""


def render_file(json_path, html_path) -> None:
    data = json.loads(Path(json_path).read_text(encoding="utf-8"))
    Path(html_path).parent.mkdir(parents=True, exist_ok=True)
    Path(html_path).write_text(render_html(data), encoding="utf-8")


def main(argv=None) -> int:
    args = list(sys.argv[1:] if argv is None else argv)
    if len(args) == 3 and args[0] in ("static", "interactive"):
        args = args[1:]
    if len(args) != 2:
        print("Usage: python -m xplus_amd.report [static|interactive] <input.json> <output.html>", file=sys.stderr)
        return 2
    render_file(args[0], args[1])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
