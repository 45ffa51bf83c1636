"""Algorithmic FLOP model of a contracted ERI (Head-Gordon-Pople Obara-Saika
count): per surviving primitive quartet

    F_prim = 30 (setup) + (12 + 3L) (Boys) + 8 N_VRR + N_acc
    N_VRR  = sum_{e=0}^{la+lb} sum_{f=0}^{lc+ld} nc(e) nc(f) (L-e-f+1) - (L+1)
    N_acc  = sum_{e=la}^{la+lb} sum_{f=lc}^{lc+ld} nc(e) nc(f)

and per surviving contracted quartet

    F_HRR  = 2 [ sum_{j=1}^{lb} sum_{i=la}^{la+lb-j} nc(i) nc(j) sum_{f=lc}^{lc+ld} nc(f)
               + sum_{j=1}^{ld} sum_{i=lc}^{lc+ld-j} nc(i) nc(j) nc(la) nc(lb) ]

with nc(l) = (l+1)(l+2)/2.  The count is lower than the McMurchie-Davidson
work the kernels do, so roofline fractions computed from it are conservative.
"""


def nc(l):
    return (l + 1) * (l + 2) // 2


def f_prim(la, lb, lc, ld):
    L = la + lb + lc + ld
    nvrr = sum(nc(e) * nc(f) * (L - e - f + 1) for e in range(la + lb + 1) for f in range(lc + ld + 1)) - (L + 1)
    nacc = sum(nc(e) * nc(f) for e in range(la, la + lb + 1) for f in range(lc, lc + ld + 1))
    return 30 + (12 + 3 * L) + 8 * nvrr + nacc


def f_hrr(la, lb, lc, ld):
    s1 = sum(nc(i) * nc(j) * sum(nc(f) for f in range(lc, lc + ld + 1))
             for j in range(1, lb + 1) for i in range(la, la + lb - j + 1))
    s2 = sum(nc(i) * nc(j) * nc(la) * nc(lb) for j in range(1, ld + 1) for i in range(lc, lc + ld - j + 1))
    return 2 * (s1 + s2)


def launch_flops(st):
    """Algorithmic flops of one class launch record (hfk.Context.class_stats)."""
    q = (st["la"], st["lb"], st["lc"], st["ld"])
    return st["prim_quartets"] * f_prim(*q) + st["shell_quartets"] * f_hrr(*q)


def f_grad_prim(la, lb, lc, ld):
    """Algorithmic flops of the nuclear-gradient contraction per surviving
    primitive quartet, as a lower bound: the derivatives with respect to
    three centres (the fourth follows from translational invariance), each
    counted at the cost of the energy class itself, f_prim(la, lb, lc, ld)
    (a derivative class has one angular momentum raised, so it costs more),
    plus 4 flops (derivative combination and the two-particle density
    product) for each of the 9 Cartesian derivatives of each component
    quartet."""
    return 3 * f_prim(la, lb, lc, ld) + 36 * nc(la) * nc(lb) * nc(lc) * nc(ld)


def grad_launch_flops(st):
    """Algorithmic flops of one class's gradient launch (the same quartets
    as its J/K launch: a class_stats record)."""
    q = (st["la"], st["lb"], st["lc"], st["ld"])
    return st["prim_quartets"] * f_grad_prim(*q)
