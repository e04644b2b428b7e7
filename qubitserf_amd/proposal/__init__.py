"""Learned information-set proposal: a Tanner-graph message-passing scorer (BF16 operands,
FP32 accumulation, HIP kernels in csrc/mp_hip.hip) whose column scores propose column
orders for the greedy information-set peel. Candidates are checked on the device (ranks of
the resulting sets, predicted BZ work) and the natural order is always a candidate, so a
proposal can only change which exact enumeration runs, never its answer."""
from .model import MPScorer, pack_batch
from .propose import choose_orders, predicted_work

__all__ = ["MPScorer", "pack_batch", "choose_orders", "predicted_work"]
