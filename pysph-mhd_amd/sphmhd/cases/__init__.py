"""Initial conditions: the benchmark workloads (workloads) and the reference's
case applications (reference-mode HCP slabs)."""
