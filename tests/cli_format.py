"""Expected report lines of the command-line driver, written from the reference application's format
strings (apps/QMIV/src/xAppQMIV.h:215-306 xMetricStat::initSuffixes / formatPerCmpMetric /
formatPerPicMetric / formatAvgMetric, apps/QMIV/src/xAppQMIV.cpp:831-935 frame lines), and helpers
that write raw sequences and run the driver."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ivpsnr_amd", "bin", "qmiv")

NAMES = ["PSNR", "WSPSNR", "IVPSNR", "SSIM", "MSSSIM", "IVSSIM"]
PER_CMP = {"PSNR", "WSPSNR", "SSIM", "MSSSIM"}
NORMALIZED = {"SSIM", "MSSSIM", "IVSSIM"}


def suffixes(name, mask, rgb):
    cmp = "-M R:G:B  " if mask and rgb else "-M Y:Cb:Cr" if mask else " R:G:B    " if rgb else " Y:Cb:Cr  "
    if name in PER_CMP:
        pic = "-M-RGB    " if mask and rgb else "-M-YCbCr  " if mask else "-RGB      " if rgb else "-YCbCr    "
    else:
        pic = "-M-RGB    " if mask and rgb else "-M        " if mask else "-RGB      " if rgb else "          "
    return cmp, pic


def frame_cmp_line(f, name, vals, mask=False, rgb=False, exact=(0, 0, 0)):
    one = "{:8.6f} " if name in NORMALIZED else "{:8.4f} "
    line = f"Frame {f:08d} " + f"{name:>8}{suffixes(name, mask, rgb)[0]} " + "".join(one.format(v) for v in vals)
    if name in ("PSNR", "WSPSNR"):
        line += "".join(t for t, e in zip((" ExactY", " ExactU", " ExactV"), exact) if e)
    return line


def frame_pic_line(f, name, val, mask=False, rgb=False):
    one = "{:8.6f} " if name in NORMALIZED else "{:8.4f} "
    return f"Frame {f:08d} " + f"{name:>8}{suffixes(name, mask, rgb)[1]} " + one.format(val) + " " * 18


def avg_lines(header, name, cmp_avg, pic_avg, mask=False, rgb=False):
    unit = "  " if name in NORMALIZED else "dB"
    one = "{:10.8f} {}  " if name in NORMALIZED else "{:10.6f} {}  "
    nf = "{:<8}{} " if header == "" else "{:>8}{} "
    sc, sp = suffixes(name, mask, rgb)
    out = []
    if name in PER_CMP:
        out.append(header + nf.format(name, sc) + "".join(one.format(v, unit) for v in cmp_avg))
    out.append(header + nf.format(name, sp) + one.format(pic_avg, unit) + " " * 18)
    return out


def pic_value(v, wavg=(4, 1, 1, 0)):
    """xMetricStat::setPerCmpMeric's weighted picture value (apps/QMIV/src/xAppQMIV.h:244-250)."""
    return (v[0] * wavg[0] + v[1] * wavg[1] + v[2] * wavg[2]) * (1.0 / sum(wavg))


def write_seq(path, frames, bd):
    with open(path, "wb") as f:
        for fr in frames:
            f.write(np.asarray(fr).astype("<u2" if bd > 8 else np.uint8).tobytes())


def run_cli(args, cwd, timeout=120):
    if not os.path.exists(CLI):
        raise FileNotFoundError("ivpsnr_amd/bin/qmiv is not built; run __graft_entry__.build() first")
    return subprocess.run([CLI] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
