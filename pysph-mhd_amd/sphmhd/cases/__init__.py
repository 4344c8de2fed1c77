"""Initial conditions: the benchmark workloads (baseline) and the reference's
case applications (reference-mode HCP slabs)."""
