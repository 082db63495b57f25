"""Fast entry points beside the qforte API: compiled UCC programs, the adjoint
gradient behind torch.autograd, and the continuous-batching serving engine."""
