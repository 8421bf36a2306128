"""Seeded synthetic tables for the benchmark workloads (no TPC-H data is available).

Every column is a counter-based stream: value(row) = f(splitmix64(seed, column, row)), so any
row range can be generated independently and the same rows come out on every machine.
Decimal columns are k/100.0 so that the literals 0.05/0.07 compare identically under the
interpreter's epsilon compare and the fused loops' exact compare.

Workloads (sizes are parameters):
  q6:  lineitem(ship_day i32 U{0..2525}, discount f64 {0.00..0.10}, quantity i32 U{1..50},
       price f64 U{900.00..105000.00}); predicate 730 <= ship_day < 1095, 0.05 <= discount
       <= 0.07, quantity < 24; sum(price*discount).
  q1:  the q6 columns plus tax f64 {0.00..0.08}; line_status = 'O' if ship_day > 1263 else
       'F'; return_flag = 'N' if ship_day + U{1..30} > 1263 else 'R'/'A' at random -- the
       TPC-H correlation that realises exactly four (flag, status) groups.
  q3:  orders(key i32 unique sparse in 1..4*n_d, order_day i32 U{0..2405}, segment i32
       U{0..4}); lineitem with 1..7 rows per order (fk), ship_day = order_day + U{1..Q3_SHIP_SPAN},
       price, discount.
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)
_M64 = (1 << 64) - 1


def _mix_seed(seed, stream):
    z = (seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03 + 0x632BE59BD9B4E019) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def rand_u64(seed, stream, n, start=0):
    """splitmix64 of (seed, stream) advanced by the row number."""
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(_mix_seed(seed, stream)) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, n, lo, hi, dtype=np.int32, start=0):
    """Integers uniform in [lo, hi]."""
    r = rand_u64(seed, stream, n, start) % np.uint64(hi - lo + 1)
    return (r.astype(np.int64) + lo).astype(dtype)


def cents(seed, stream, n, lo, hi, start=0):
    """k/100.0 for k uniform in [lo, hi]."""
    return uniform(seed, stream, n, lo, hi, np.int64, start).astype(np.float64) / 100.0


Q6_LO, Q6_HI = 730, 1095
Q1_CUT = 1263
Q1_SHIP_MAX = 2434
Q3_DAY = 1170


def gen_q6(n, seed=1000):
    return dict(
        ship_day=uniform(seed, 1, n, 0, 2525),
        discount=cents(seed, 2, n, 0, 10),
        quantity=uniform(seed, 3, n, 1, 50),
        price=cents(seed, 4, n, 90000, 10500000),
    )


def gen_q1(n, seed=1001):
    t = gen_q6(n, seed)
    t["tax"] = cents(seed, 5, n, 0, 8)
    ship = t["ship_day"]
    receipt = ship + uniform(seed, 6, n, 1, 30)
    ra = np.where(uniform(seed, 7, n, 0, 1) == 1, ord("R"), ord("A"))
    t["return_flag"] = np.where(receipt > Q1_CUT, ord("N"), ra).astype(np.uint8)
    t["line_status"] = np.where(ship > Q1_CUT, ord("O"), ord("F")).astype(np.uint8)
    return t


# Days from an order to the shipping of its lineitems, U{1..Q3_SHIP_SPAN}.  TPC-H's span of 121
# days realises about 11.4K result groups at 1.5M orders (30K joined rows); the join + group
# workload is stated at about 30K groups, which a span of 318 days gives (30,091 groups, 79,648
# joined rows at 1.5M orders, seed 1002).
Q3_SHIP_SPAN = 318


def _affine_perm(n, seed):
    """A bijection of [0, n): i -> (a*i + b) mod n with gcd(a, n) = 1."""
    a = int(_mix_seed(seed, 99) % max(n, 1)) | 1
    while np.gcd(a, n) != 1:
        a += 2
    b = int(_mix_seed(seed, 98) % max(n, 1))
    i = np.arange(n, dtype=np.int64)
    return (a * i + b) % n


def gen_q3(n_dim, seed=1002):
    perm = _affine_perm(n_dim, seed)
    key = (4 * perm + 1 + uniform(seed, 11, n_dim, 0, 3, np.int64)).astype(np.int32)
    order_day = uniform(seed, 12, n_dim, 0, 2405)
    segment = uniform(seed, 13, n_dim, 0, 4)
    per = uniform(seed, 14, n_dim, 1, 7, np.int64)
    owner = np.repeat(np.arange(n_dim, dtype=np.int64), per)
    nf = owner.shape[0]
    fact = dict(
        fk=key[owner],
        ship_day=(order_day[owner] + uniform(seed, 15, nf, 1, Q3_SHIP_SPAN)).astype(np.int32),
        price=cents(seed, 16, nf, 90000, 10500000),
        discount=cents(seed, 17, nf, 0, 10),
    )
    dim = dict(key=key, order_day=order_day, segment=segment)
    return dim, fact


SERVE_FACT = ("fk", "ship_day", "price", "discount", "quantity", "tax", "return_flag", "line_status")
SERVE_DIM = ("key", "order_day", "segment")


def gen_serve(n_dim, seed=1004):
    """Orders and one lineitem table carrying the columns of q1, q3 and q6 (the serving
    tables); the host twin of the device generator hpaGenServeTables."""
    perm = _affine_perm(n_dim, seed)
    key = (4 * perm + 1 + uniform(seed, 11, n_dim, 0, 3, np.int64)).astype(np.int32)
    order_day = uniform(seed, 12, n_dim, 0, 2405)
    segment = uniform(seed, 13, n_dim, 0, 4)
    per = uniform(seed, 14, n_dim, 1, 7, np.int64)
    owner = np.repeat(np.arange(n_dim, dtype=np.int64), per)
    nf = owner.shape[0]
    ship = (order_day[owner] + uniform(seed, 15, nf, 1, Q3_SHIP_SPAN)).astype(np.int32)
    receipt = ship + uniform(seed, 20, nf, 1, 30)
    ra = np.where(uniform(seed, 21, nf, 0, 1) == 1, ord("R"), ord("A"))
    fact = dict(
        fk=key[owner], ship_day=ship,
        price=cents(seed, 16, nf, 90000, 10500000), discount=cents(seed, 17, nf, 0, 10),
        quantity=uniform(seed, 18, nf, 1, 50), tax=cents(seed, 19, nf, 0, 8),
        return_flag=np.where(receipt > Q1_CUT, ord("N"), ra).astype(np.uint8),
        line_status=np.where(ship > Q1_CUT, ord("O"), ord("F")).astype(np.uint8),
    )
    return dict(key=key, order_day=order_day, segment=segment), fact


def gen_batch_query(q, n, card, seed=1003):
    """One batched filter->compress->group->sum query: keys uniform over `card` values."""
    s = seed * 1000 + q
    return dict(
        key=uniform(s, 21, n, 0, card - 1),
        val=cents(s, 22, n, 0, 10000000),
        flt=uniform(s, 23, n, 0, 99),
    )


def batch_shapes(nq=64, seed=1003, nmin=1e4, nmax=1e7, gmin=4, gmax=1e6):
    """Row counts and group cardinalities, log-uniform, one pair per query."""
    u = rand_u64(seed, 31, 2 * nq).astype(np.float64) / 2.0**64
    n = np.exp(np.log(nmin) + u[:nq] * (np.log(nmax) - np.log(nmin))).astype(np.int64)
    g = np.exp(np.log(gmin) + u[nq:] * (np.log(gmax) - np.log(gmin))).astype(np.int64)
    thr = uniform(seed, 32, nq, 20, 80)
    return n, np.maximum(g, 1), thr
