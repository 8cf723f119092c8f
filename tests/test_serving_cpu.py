"""Host-side serving logic on the CPU: HBM-budget admission with an injected
budget (reject / queue / admit) and an injected creation error (the request
fails alone, the others are admitted)."""

from jamma_amd.serve.server import AdmissionQueue


def test_admission_first_fit_queue_and_release():
    q = AdmissionQueue(budget=100)
    q.queue = ["a60", "b50", "c30", "d10"]
    cost = lambda r: int(r[1:])  # noqa: E731
    adm, failed = q.admit(cost, lambda r: r.upper())
    assert [x[0] for x in adm] == ["a60", "c30", "d10"] and not failed
    assert q.queue == ["b50"] and q.used == 100
    q.release(60)
    adm, _ = q.admit(cost, lambda r: r)
    assert [x[0] for x in adm] == ["b50"] and q.used == 90 and q.queue == []


def test_admission_rejects_never_fitting_and_isolates_errors():
    q = AdmissionQueue(budget=50)
    assert not q.fits_ever(51)

    def create(r):
        if r == "bad":
            raise RuntimeError("jm_lmm_null failed (rc=-1): injected")
        return r

    q.queue = ["x10", "bad", "y20"]
    cost = lambda r: 5 if r == "bad" else int(r[1:])  # noqa: E731
    adm, failed = q.admit(cost, create)
    assert [x[0] for x in adm] == ["x10", "y20"]
    assert failed[0][0] == "bad" and "injected" in str(failed[0][1])
    assert q.used == 30  # the failed request holds no budget
