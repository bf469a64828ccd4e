"""Learned warm start for the batched TV solver (no counterpart in the reference)."""
from opymize_amd.warmstart.cnn import ConvNet, WarmStart, encode_inputs, encode_targets

__all__ = ["ConvNet", "WarmStart", "encode_inputs", "encode_targets"]
