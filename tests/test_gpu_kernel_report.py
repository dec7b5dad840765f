"""The streaming-pass report (research_amd/kernel_report.py, also part of every bench line)
runs end to end and reports every pass with a positive time and bandwidth."""
import pytest

from research_amd import kernel_report

pytestmark = pytest.mark.gpu


def test_kernel_report_reports_every_pass(ctx):
    r = kernel_report.report(ctx, 256, 144, reps=2, em_samples=4096)
    assert set(r["passes"]) == set(kernel_report.BYTES)
    for name, v in r["passes"].items():
        assert v["ms"] > 0.0 and v["gbs"] > 0.0, name
    assert r["copy_gbs"] > 0.0 and r["em"]["samples_per_step"] == 4096
