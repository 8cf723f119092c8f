"""Host-side serving logic on the CPU: HBM-budget admission in arrival order with
first fit (reject / queue / admit / release) and aging (a request overtaken
max_skips times stops further overtaking until it fits)."""

from jamma_amd.serve.server import AdmissionQueue


def _cost(r):
    return int(r[1:])


def test_admission_first_fit_queue_and_release():
    q = AdmissionQueue(budget=100)
    q.queue = ["a60", "b50", "c30", "d10"]
    adm = q.admit(_cost)
    assert [x[0] for x in adm] == ["a60", "c30", "d10"] and [x[1] for x in adm] == [60, 30, 10]
    assert q.queue == ["b50"] and q.used == 100
    q.release(60)
    adm = q.admit(_cost)
    assert [x[0] for x in adm] == ["b50"] and q.used == 90 and q.queue == []


def test_admission_rejects_never_fitting():
    q = AdmissionQueue(budget=50)
    assert not q.fits_ever(51) and q.fits_ever(50)


def test_aging_stops_overtaking():
    """A big request overtaken max_skips times blocks the small ones behind it until
    budget frees up for it; then it is admitted before them."""
    q = AdmissionQueue(budget=100, max_skips=3)
    q.used = 60  # someone else holds 60
    big = "B80"
    q.queue = [big]
    for _ in range(3):  # three rounds of small requests overtake it
        q.queue.append("x10")
        adm = q.admit(_cost)
        assert [x[0] for x in adm] == ["x10"]
        q.release(10)
    assert q.skips(big) == 3
    q.queue.append("y10")
    assert q.admit(_cost) == []  # aged: y10 may not pass B80 although it fits
    assert q.queue == [big, "y10"]
    q.release(60)  # the holder finishes
    adm = q.admit(_cost)
    assert [x[0] for x in adm] == [big, "y10"] and q.used == 90
    assert q.skips(big) == 0
