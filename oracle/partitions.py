"""Edge basis enumeration -- oracle restatement (test infrastructure only).

Restates the reference's basis construction, modules/partitions/integer_partitions.cu:
  * accelAsc enumeration of the partitions of dL (Kelleher) ........ :10-47
  * drop partitions with more than N parts ......................... :104-117 (p.size() <= cutoff)
  * compress to (part, multiplicity) pairs in increasing part order . :118-131
  * stable bubble sort by increasing multiplicity .................. :63-77
  * flatten over the requested sectors, in the order given ......... :199-226
dL = 0 yields the single partition (0 x 1), whose polynomial is taken as 1.
"""


def accel_asc(n):
    """All partitions of n as ascending part lists, in accelAsc order.

    integer_partitions.cu:10-47 (a[] has n+1 slots; the inclusive `extract(a,0,k+1)`
    is a[:k+2]).
    """
    a = [0] * (n + 2)
    k = 1
    y = n - 1
    out = []
    while k != 0:
        x = a[k - 1] + 1
        k -= 1
        while 2 * x <= y:
            a[k] = x
            y -= x
            k += 1
        l = k + 1
        while x <= y:
            a[k] = x
            a[l] = y
            out.append(list(a[:k + 2]))
            x += 1
            y -= 1
        a[k] = x + y
        y = x + y - 1
        out.append(list(a[:k + 1]))
    return out


def compress(parts):
    """(part, multiplicity) pairs, increasing part, then stable sort by multiplicity
    (integer_partitions.cu:118-133 and :63-77)."""
    lo, hi = min(parts), max(parts)
    pairs = [(v, parts.count(v)) for v in range(lo, hi + 1) if parts.count(v)]
    # bubble sort on repetitions, swapping only when strictly greater (stable)
    pairs = list(pairs)
    n = len(pairs)
    for i in range(n):
        for j in range(n - i - 1):
            if pairs[j][1] > pairs[j + 1][1]:
                pairs[j], pairs[j + 1] = pairs[j + 1], pairs[j]
    return pairs


def sector_basis(n_particles, dl):
    """Compressed partitions of dl with at most n_particles parts."""
    return [compress(p) for p in accel_asc(dl) if len(p) <= n_particles]


class Basis:
    """Flattened basis over a list of sectors."""

    def __init__(self, n_particles, sectors):
        self.N = int(n_particles)
        self.sectors = [int(s) for s in sectors]
        self.states = []      # list of (dL, compressed pairs)
        self.sector_offsets = [0]
        for dl in self.sectors:
            sb = sector_basis(self.N, dl)
            self.states += [(dl, p) for p in sb]
            self.sector_offsets.append(len(self.states))
        self.D = len(self.states)
        self.K = max([k for _, p in self.states for k, _ in p] + [0])
        self.longest = max([len(p) for _, p in self.states] + [0])

    def sector_dims(self):
        o = self.sector_offsets
        return [o[i + 1] - o[i] for i in range(len(self.sectors))]

    def factor_arrays(self):
        """fo (D+1), fk, fr arrays for the oracle library."""
        fo, fk, fr = [0], [], []
        for _, p in self.states:
            for k, r in p:
                fk.append(k)
                fr.append(r)
            fo.append(len(fk))
        return fo, fk, fr
