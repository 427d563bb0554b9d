"""MI355X-native fused RAW develop: the reference's `flashraw-gpu` develop entry points over
hand-written gfx950 HIP kernels (flashraw_amd/lib/libflashraw_hip.so)."""
from .gpu import (Batch, BlackLevel, Cancelled, CancelToken, CfaColor, CfaPattern, CfaDenoiseSpec, DenoiseSpec, DevelopOutput, DnPrecision, DevelopSpec, EDGE_CANON, EDGE_CANON_XTRANS, EDGE_HALO, EdgeSpec, FlashRawError,
                  FrameInfo, GpuContext, GpuCurve, GpuFrame, GpuImage, GpuOp, GpuRun, GpuTimings, GpuTone, GpuUnavailable,
                  InvalidArgument, JobResult, PackOrder, LimitExceeded, NotAvailable, Orientation, OutOfMemory, Rect, Server,
                  NoiseReduction, PlaneSpec, WhiteLevel, pinned_empty,
                  ahd_spec, available, context, denoise_spec, edge_plan, edge_spec, estimate_beta, lib, nr_model_fit, nr_model_from_stats,
                  srgb8_table)

__all__ = [
    "Batch", "BlackLevel", "Cancelled", "CancelToken", "CfaColor", "CfaPattern", "CfaDenoiseSpec", "DenoiseSpec", "DevelopOutput", "DnPrecision", "DevelopSpec", "EDGE_CANON", "EDGE_CANON_XTRANS", "EDGE_HALO", "EdgeSpec", "FlashRawError",
    "FrameInfo", "GpuContext", "GpuCurve", "GpuFrame", "GpuImage", "GpuOp", "GpuRun", "GpuTimings", "GpuTone", "GpuUnavailable",
    "InvalidArgument", "JobResult", "PackOrder", "LimitExceeded", "NotAvailable", "Orientation", "OutOfMemory", "Rect", "Server",
    "NoiseReduction", "PlaneSpec", "WhiteLevel", "pinned_empty",
    "ahd_spec", "available", "context", "denoise_spec", "edge_plan", "edge_spec", "estimate_beta", "lib",
    "nr_model_fit", "nr_model_from_stats", "srgb8_table",
]
