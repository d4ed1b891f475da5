"""Reference-compatible operator module (``warpgbm.cuda.node_kernel``)."""
