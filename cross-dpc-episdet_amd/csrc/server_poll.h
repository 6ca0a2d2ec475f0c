// server_poll.h -- the poll protocol of the server's request table.
//
// No reference counterpart (the reference has no request loop).  Header-only
// and free of HIP, so that the host test build runs it under AddressSanitizer
// with real threads.
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <mutex>

#include "../../include/episdet.h"

namespace epi {

// Collects request `id` from `all` (id -> std::unique_ptr<R>; R has
// `bool done, polling` and `int status`), with the table's mutex held by
// `lk`; `cv` is notified whenever a request completes.
//   unknown id (never submitted, or already collected)  EPI_ENOENT
//   another thread is polling the same id                EPI_EBUSY
//   not done within timeout_ms (negative: no limit)      EPI_EAGAIN
//   done: take(*r) copies the result out, the entry is erased, and the
//   request's status is returned.
// A poller marks its request while it waits (the lock is released meanwhile),
// so no second poller can collect and free it under the first.  `inside`
// counts the pollers in this function, for an owner that must not free the
// table while one still waits (it waits on `cv` for inside == 0).
template <class Map, class Take>
int poll_request(Map &all, uint64_t id, std::unique_lock<std::mutex> &lk,
                 std::condition_variable &cv, int32_t timeout_ms, uint32_t &inside, Take take)
{
    auto it = all.find(id);
    if (it == all.end())
        return EPI_ENOENT;
    auto *r = it->second.get();
    if (r->polling)
        return EPI_EBUSY;
    r->polling = true;
    ++inside;
    auto ready = [r] { return r->done; };
    bool ok = true;
    if (timeout_ms < 0)
        cv.wait(lk, ready);
    else
        ok = cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    --inside;
    int st = EPI_EAGAIN;
    if (ok) {
        st = r->status;
        take(*r);
        all.erase(id);  // by key: other submissions may have rehashed the table
    } else {
        r->polling = false;
    }
    cv.notify_all();
    return st;
}

}  // namespace epi
