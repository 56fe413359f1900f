"""FEM-FEM penalty contact: the batched work list of a loop's FEM-FEM pairs.

Follows the FlexFlex part of MixedContact._build_batched_contacts
(flatworld/mixedcontact.py:1857-1957) and the FEM spatial hash's sizing
(flatworld/femspringmanager.py:520-556; flatworld/spatialmanager.py:508-565):

* the group's bodies are every FEM body of the loop, in the manager's order; the
  hash holds every body's boundary edges, body by body;
* for each FEM-FEM pair, in contact order: the boundary nodes of ``domain2``
  checked against ``domain1``'s edges, then ``domain1``'s against ``domain2``'s,
  with the pair's float32 penalty and the larger of the two frictions;
* hash limits: 100000 cells, 512 query results, and 9 pairs per element slot of
  ``MAX_BOUNDARY_ELEMENTS = max(boundary elements + 1024, MAX_NODES)``;
* a tied contact (``contact.tied``, set by ExplicitLoop's ``ties``) flags every
  work item of its pair; tied items are resolved once to their nearest edge and
  then held there by a penalty spring instead of sliding
  (flatworld/mixedcontact.py:687-811, 1871-1932).

The contact itself runs on the device (flatworld_amd/csrc/fw_flex.hip).
"""

from dataclasses import dataclass

import numpy as np

# simulated time between hash rebuilds (flatworld/femspringmanager.py:524, 1035-1039)
REBUILD_INTERVAL = 1.0 / 500.0
MAX_CELLS = 100000
MAX_QUERY = 512


@dataclass
class FlexGroupSpec:
    """Flat arrays of one contact group; node ids are body-local."""

    body_seg_off: np.ndarray  # int32[B + 1]
    seg_node: np.ndarray  # int32[S, 2]
    body_bnodes_off: np.ndarray  # int32[B + 1]
    body_bnodes: np.ndarray  # int32[sum Nb]
    item_body: np.ndarray  # int32[W]
    item_node: np.ndarray
    item_target: np.ndarray
    item_pair: np.ndarray
    item_penalty: np.ndarray  # float32[W]
    item_friction: np.ndarray
    pair_body: np.ndarray  # int32[P, 2]
    gated: bool
    item_tied: np.ndarray = None  # int32[W], 1 where the item's pair is tied (None: none)
    max_cells: int = MAX_CELLS
    max_query: int = MAX_QUERY
    max_pairs: int = 9 * 4096
    rebuild_interval: float = REBUILD_INTERVAL

    @property
    def n_bodies(self):
        return int(self.body_seg_off.shape[0] - 1)

    @property
    def n_segs(self):
        return int(self.seg_node.shape[0])

    @property
    def n_items(self):
        return int(self.item_body.shape[0])

    @property
    def n_pairs(self):
        return int(self.pair_body.shape[0])

    def block_bytes_bound(self):
        """An upper bound of the arena bytes fw_flex_add takes for this group: uploaded
        tables, hash arrays and per-substep scratch, each padded to 256 bytes."""
        W, S, B = self.n_items, self.n_segs, self.n_bodies
        NS = int(self.body_bnodes.shape[0])
        return (168 * W + 60 * S + 12 * NS + 8 * self.n_pairs + 24 * B + 16 * self.max_cells + 8 * self.max_pairs
                + 48 * 256)

    def tied_flags(self):
        return np.zeros(self.n_items, np.int32) if self.item_tied is None else np.asarray(self.item_tied, np.int32)

    @property
    def n_tied(self):
        return int(self.tied_flags().sum())

    def in_pair(self, b):
        return bool(np.any(self.pair_body == b))

    def flex_cap(self, b):
        """Most contributions body b can receive in one substep: one per own work
        item, two per work item checked against its edges."""
        return int(np.sum(self.item_body == b) + 2 * np.sum(self.item_target == b))


def build_flex_group(fem_domains, flex_contacts, gated, count_nodes):
    """Work list of ``flex_contacts`` (ContactFlexFlex, in contact order) over the
    loop's FEM bodies ``fem_domains`` (manager order).  ``count_nodes`` is the
    manager's node count (FEM and spring-mass), which sizes the hash's pair buffer."""
    idx = {id(d): k for k, d in enumerate(fem_domains)}
    seg_off, segs, bn_off, bns = [0], [], [0], []
    n_be = 0
    for d in fem_domains:
        m = d.mesh
        be = np.asarray(m.boundaryElements, dtype=np.int32).reshape(-1, 2) if m.numBoundElements else np.zeros((0, 2), np.int32)
        segs.append(be)
        seg_off.append(seg_off[-1] + be.shape[0])
        n_be += be.shape[0]
        bn = np.asarray(m.boundaryNodes, dtype=np.int32) if m.numBoundNodes else np.zeros(0, np.int32)
        bns.append(bn)
        bn_off.append(bn_off[-1] + bn.shape[0])
    ib, inode, it, ip, ipen, ifr, itied, pairs = [], [], [], [], [], [], [], []
    for q, c in enumerate(flex_contacts):
        d1, d2 = idx[id(c.domain1)], idx[id(c.domain2)]
        pen = np.float32(float(c.penalty))
        fric = np.float32(max(float(c.domain1.friction), float(c.domain2.friction)))
        tied = 1 if getattr(c, "tied", False) else 0
        pairs.append((d1, d2))
        for src, dst in ((d2, d1), (d1, d2)):
            n = bns[src].shape[0]
            ib.append(np.full(n, src, np.int32))
            inode.append(bns[src])
            it.append(np.full(n, dst, np.int32))
            ip.append(np.full(n, q, np.int32))
            ipen.append(np.full(n, pen, np.float32))
            ifr.append(np.full(n, fric, np.float32))
            itied.append(np.full(n, tied, np.int32))
    cat = lambda a, t: np.ascontiguousarray(np.concatenate(a).astype(t)) if a else np.zeros(0, t)  # noqa: E731
    max_nodes = max(count_nodes + 1024, 2048)
    max_elements = max(n_be + 1024, max_nodes)
    return FlexGroupSpec(
        body_seg_off=np.asarray(seg_off, np.int32),
        seg_node=np.ascontiguousarray(np.concatenate(segs).astype(np.int32)) if segs else np.zeros((0, 2), np.int32),
        body_bnodes_off=np.asarray(bn_off, np.int32),
        body_bnodes=cat(bns, np.int32),
        item_body=cat(ib, np.int32),
        item_node=cat(inode, np.int32),
        item_target=cat(it, np.int32),
        item_pair=cat(ip, np.int32),
        item_penalty=cat(ipen, np.float32),
        item_friction=cat(ifr, np.float32),
        pair_body=np.asarray(pairs, np.int32).reshape(-1, 2),
        gated=bool(gated),
        item_tied=cat(itied, np.int32),
        max_pairs=9 * max_elements,
    )
