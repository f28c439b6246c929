"""Hand-traced OptimizeBandwidth results (fem/src/BandwidthOptimize.src:108-355)
for two tiny meshes whose scrambled numbering makes the reordering pay, so
it is accepted.  Each expected Perm was worked out by following the
reference routine line by line; the steps are recorded beside the data.

Facts of the routine used below:
- ListMatrix rows are sorted and include the node itself (List_GetMatrixIndex,
  ListMatrix.src:270-320: every element inserts all its (i, j) pairs), and
  Degree counts the entries.  With all nodes active Perm = identity before
  the optimisation (ElementUtils.src:765-773).
- Half bandwidth = max |i - k| over list entries, plus 1 (:136, :243-244).
- Start node: the first node of smallest Degree (:156-165).  The NewRoot
  loop (:172-200) then looks at the nodes on the last level of the
  Levelize DFS for a Degree strictly below the start node's; the start node
  already has the smallest Degree of all, so that search never finds one
  and the start node (and the `StartNode = j` line at :197) never changes.
- Main loop (:205-225): PermLocal(1) = start; node PermLocal(i)'s list is
  scanned in order and every node not yet numbered is appended (Renumber,
  :266-286); when the queue runs dry (PermLocal(i) = 0) the smallest
  unnumbered node starts the next component.
- Reversal (:227-240): Perm(node) = LocalNodes - (its place in PermLocal) + 1.
- Rejected only when the bandwidth before is smaller than after (:251).
"""
import numpy as np

CASES = {
    # 2 x 2 Quad4 square, 9 nodes.  Grid positions (x, y) carry the labels
    #   (0,2)=3 (1,2)=8 (2,2)=5
    #   (0,1)=4 (1,1)=1 (2,1)=6
    #   (0,0)=7 (1,0)=2 (2,0)=9
    # Lists: 1:{1..9} deg 9; 2:{1,2,4,6,7,9} 6; 3:{1,3,4,8} 4; 4:{1,2,3,4,7,8} 6;
    #        5:{1,5,6,8} 4; 6:{1,2,5,6,8,9} 6; 7:{1,2,4,7} 4; 8:{1,3,4,5,6,8} 6;
    #        9:{1,2,6,9} 4.
    # Before: node 1 lists node 9 -> |1-9| = 8 -> 9.
    # Start: degrees 9,6,4,... -> node 3 (first of degree 4).
    # Queue: 3 | from L3: 1, 4, 8 | from L1: 2, 5, 6, 7, 9 ->
    #   PermLocal = 3 1 4 8 2 5 6 7 9.
    # Reversed: 3->9, 1->8, 4->7, 8->6, 2->5, 5->4, 6->3, 7->2, 9->1.
    # After: node 1 (now 8) meets node 9 (now 1): 7 -> 8 <= 9, accepted.
    "quad4_2x2_scrambled": {
        "nnodes": 9,
        "elements": [[7, 2, 1, 4], [2, 9, 6, 1], [4, 1, 8, 3], [1, 6, 5, 8]],
        "perm": [8, 5, 9, 7, 4, 3, 2, 6, 1],
        "bw_before": 9, "bw_after": 8, "accepted": True,
    },
    # Two separate chains of 2-node elements: 5-2-6 and 4-1-3.
    # Lists: 1:{1,3,4} deg 3; 2:{2,5,6} 3; 3:{1,3} 2; 4:{1,4} 2; 5:{2,5} 2; 6:{2,6} 2.
    # Before: node 2 lists node 6 -> 4 -> 5.
    # Start: node 3 (first of degree 2; nodes 4, 5, 6 tie and lose).
    # Queue: 3 | L3: 1 | L1: 4 | L4: nothing | queue dry at place 4: the
    #   smallest unnumbered node, 2 | L2: 5, 6 -> PermLocal = 3 1 4 2 5 6.
    # Reversed: 3->6, 1->5, 4->4, 2->3, 5->2, 6->1.
    # After: the largest gap is 2 (node 2 = 3 with node 6 = 1) -> 3 <= 5, accepted.
    "two_chains_ties": {
        "nnodes": 6,
        "elements": [[5, 2], [2, 6], [4, 1], [1, 3]],
        "perm": [5, 3, 6, 4, 2, 1],
        "bw_before": 5, "bw_after": 3, "accepted": True,
    },
}


def case(name):
    c = CASES[name]
    return (np.asarray(c["elements"], np.int32), c["nnodes"], np.asarray(c["perm"], np.int32),
            c["bw_before"], c["bw_after"], c["accepted"])
