"""Oracle: FP64 CPU restatement of the reference pipeline.  Test infrastructure
only -- imported by tests/, by __graft_entry__.smoke() and by bench.py's
CPU-baseline leg as the checker, never by the product."""
