"""Autotuner of the FFT radices and tiles (dnsbox_amd/csrc/tuned_cfg.hpp); see autotune.py."""
