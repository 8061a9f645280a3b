"""FP64 CPU oracle for the Gaussian kernel-sum engine — test infrastructure only."""
