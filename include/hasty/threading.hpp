// Host thread pool with per-thread storage: the reference's threading module
// (cpp/corelib/lib/threading/threading.cppm:15-395), C++17, no device code.
//
//   storage                        named, typed shared_ptr slots owned by one
//                                  worker thread (a device index, plans, device
//                                  copies, ...)
//   storage_thread_pool            one worker per storage; a task is routed to a
//                                  worker whose storage has every name of
//                                  must_have_all and one of must_have_any, and
//                                  preferably every name of good_to_have
//   storage_thread_pool_interface  a pool plus fixed requirement sets
//
// Scheduling follows the reference (threading.cppm:130-141, 276-323): a worker
// takes the first eligible task whose good_to_have names it holds in full, and
// otherwise the eligible task with the most of them once that task has waited
// task_optimality_wait_time milliseconds.  A task's exception reaches the
// caller through its future.
//
// Three differences, all defects of the reference: a task with no
// good_to_have names starts at once, as the reference's documentation says
// (:136-138; its code makes such a task wait the full optimality time, :295);
// clear_from_storages wakes up when the last task finishes (the reference
// waits on a count that no worker signals, :231 vs 321); and the pool
// takes tasks again once the clear is done (the reference sets a pause flag
// that nothing reads or resets, :225).
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <future>
#include <list>
#include <memory>
#include <mutex>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <typeindex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace hasty {
namespace threading {

// the reference's vset<std::string> (an ordered set of names)
using name_set = std::set<std::string>;

class storage {
    struct item {
        std::shared_ptr<void> ptr;
        std::type_index type;
    };
    std::unordered_map<std::string, item> _storage;

    template <typename T> [[noreturn]] static void wrong_type(const std::string& name) {
        throw std::runtime_error("Storage does not contain key: " + name + " with type: " + typeid(T).name());
    }
    [[noreturn]] static void missing(const std::string& name) {
        throw std::runtime_error("Storage does not contain key: " + name);
    }

public:
    // nullopt when the name is absent or holds another type (threading.cppm:27-37)
    template <typename T> std::optional<std::reference_wrapper<T>> get_ref(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end() || it->second.type != std::type_index(typeid(T))) return std::nullopt;
        return std::ref(*static_cast<T*>(it->second.ptr.get()));
    }
    template <typename T> T& get_ref_throw(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end()) missing(name);
        if (it->second.type != std::type_index(typeid(T))) wrong_type<T>(name);
        return *static_cast<T*>(it->second.ptr.get());
    }
    template <typename T> std::shared_ptr<T> get_ptr(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end() || it->second.type != std::type_index(typeid(T))) return nullptr;
        return std::static_pointer_cast<T>(it->second.ptr);
    }
    template <typename T> std::shared_ptr<T> get_ptr_throw(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end()) missing(name);
        if (it->second.type != std::type_index(typeid(T))) wrong_type<T>(name);
        return std::static_pointer_cast<T>(it->second.ptr);
    }
    // keeps an existing entry of the same name, as the reference's insert does
    template <typename T> void add(const std::string& name, std::shared_ptr<T> ptr) {
        _storage.emplace(name, item{std::static_pointer_cast<void>(std::move(ptr)), std::type_index(typeid(T))});
    }
    template <typename T> void clear(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end()) return;
        if (it->second.type != std::type_index(typeid(T))) wrong_type<T>(name);
        _storage.erase(it);
    }
    template <typename T> void clear_throw(const std::string& name) {
        auto it = _storage.find(name);
        if (it == _storage.end()) missing(name);
        if (it->second.type != std::type_index(typeid(T))) wrong_type<T>(name);
        _storage.erase(it);
    }
    void clear(const std::string& name) { _storage.erase(name); }
    bool exist(const std::string& name) const { return _storage.count(name) != 0; }
    name_set names() const {
        name_set r;
        for (const auto& kv : _storage) r.insert(kv.first);
        return r;
    }
};

class storage_thread_pool {
public:
    explicit storage_thread_pool(std::vector<storage>&& storages, int task_optimality_wait_time = 100)
        : _storages(std::move(storages)), _wait_ms(task_optimality_wait_time) {
        _threads.reserve(_storages.size());
        try {
            for (size_t i = 0; i < _storages.size(); ++i) _threads.emplace_back([this, i] { work(i); });
        } catch (...) {
            stop();
            throw;
        }
    }
    storage_thread_pool(const storage_thread_pool&) = delete;
    storage_thread_pool& operator=(const storage_thread_pool&) = delete;
    ~storage_thread_pool() { stop(); }

    template <class F, class... Args>
    std::future<std::invoke_result_t<F, storage&, Args...>> enqueue(F&& f, name_set must_have_all_names,
                                                                   name_set must_have_any_names,
                                                                   name_set good_to_have_names, Args&&... args) {
        using R = std::invoke_result_t<F, storage&, Args...>;
        auto task = std::make_shared<std::packaged_task<R(storage&)>>(
            [f = std::forward<F>(f), tup = std::make_tuple(std::forward<Args>(args)...)](storage& s) mutable -> R {
                return std::apply([&](auto&... a) -> R { return std::invoke(f, s, a...); }, tup);
            });
        std::future<R> res = task->get_future();
        {
            std::lock_guard<std::mutex> lock(_mutex);
            if (_stop) throw std::runtime_error("enqueue on a stopped storage_thread_pool");
            const bool held = _paused > 0;
            _work.push_back(work_item{[task](storage& s) { (*task)(s); }, std::move(must_have_all_names),
                                      std::move(must_have_any_names), std::move(good_to_have_names),
                                      std::chrono::steady_clock::now(), held});
            if (held) ++_held;
            else ++_queued, ++_pending;
        }
        _cv.notify_all();
        return res;
    }

    size_t nthreads() const { return _storages.size(); }
    // the workers' storages; read them while no task runs (tasks change them)
    const std::vector<storage>& storages() const { return _storages; }

    // Waits for every task enqueued before the call, then removes
    // names_to_clear from each storage that holds all of storage_match_all and
    // one of storage_match_any (any storage when that set is empty).  Tasks
    // enqueued meanwhile start once the clear is done.  Not to be called from
    // inside a task: it would wait for itself.
    void clear_from_storages(const name_set& storage_match_all, const name_set& storage_match_any,
                             const name_set& names_to_clear) {
        std::unique_lock<std::mutex> lock(_mutex);
        ++_paused;
        _idle.wait(lock, [this] { return _pending == 0; });
        for (auto& s : _storages) {
            const bool all = std::all_of(storage_match_all.begin(), storage_match_all.end(),
                                         [&](const std::string& n) { return s.exist(n); });
            const bool any = storage_match_any.empty() ||
                             std::any_of(storage_match_any.begin(), storage_match_any.end(),
                                         [&](const std::string& n) { return s.exist(n); });
            if (all && any)
                for (const auto& n : names_to_clear) s.clear(n);
        }
        if (--_paused == 0 && _held) {  // release the tasks held back by the clear
            for (auto& w : _work) w.held = false;
            _queued += _held;
            _pending += _held;
            _held = 0;
        }
        lock.unlock();
        _cv.notify_all();
    }

private:
    struct work_item {
        std::function<void(storage&)> func;
        name_set must_all, must_any, good_to_have;
        std::chrono::steady_clock::time_point enqueue_time;
        bool held;  // enqueued during a clear: not started before it ends
    };

    void stop() {
        {
            std::lock_guard<std::mutex> lock(_mutex);
            _stop = true;
        }
        _cv.notify_all();
        for (auto& t : _threads)
            if (t.joinable()) t.join();
    }

    // the best task for this storage, or _work.end() (threading.cppm:276-311)
    std::list<work_item>::iterator pick(const storage& store, std::chrono::steady_clock::time_point now) {
        auto best = _work.end();
        int best_good = -1;
        for (auto it = _work.begin(); it != _work.end(); ++it) {
            if (it->held) continue;
            if (!std::all_of(it->must_all.begin(), it->must_all.end(),
                             [&](const std::string& n) { return store.exist(n); }))
                continue;
            if (!it->must_any.empty() && std::none_of(it->must_any.begin(), it->must_any.end(),
                                                      [&](const std::string& n) { return store.exist(n); }))
                continue;
            int good = 0;
            for (const auto& n : it->good_to_have) good += store.exist(n) ? 1 : 0;
            if (good == int(it->good_to_have.size())) return it;  // perfect match (or nothing asked)
            if (good > best_good) {
                best_good = good;
                best = it;
            }
        }
        if (best != _work.end() && now - best->enqueue_time <= std::chrono::milliseconds(_wait_ms))
            return _work.end();  // let a better-placed worker take it for a while
        return best;
    }

    void work(size_t index) {
        storage& store = _storages[index];
        std::unique_lock<std::mutex> lock(_mutex);
        for (;;) {
            if (_stop) return;
            auto it = pick(store, std::chrono::steady_clock::now());
            if (it == _work.end()) {
                // nothing for this worker now: sleep until new work or a clear
                // ending, polling while a task waits out its optimality wait
                if (_queued == 0) _cv.wait(lock);
                else _cv.wait_for(lock, std::chrono::milliseconds(2));
                continue;
            }
            auto func = std::move(it->func);
            _work.erase(it);
            --_queued;
            lock.unlock();
            func(store);  // a packaged_task: exceptions go to the future
            lock.lock();
            if (--_pending == 0) _idle.notify_all();
        }
    }

    std::vector<storage> _storages;
    std::vector<std::thread> _threads;
    int _wait_ms;
    std::mutex _mutex;
    std::condition_variable _cv, _idle;
    std::list<work_item> _work;
    size_t _queued = 0;   // in _work and not held
    size_t _pending = 0;  // _queued + running
    size_t _held = 0;     // in _work and held
    int _paused = 0;      // clears in progress
    bool _stop = false;
};

class storage_thread_pool_interface {
public:
    storage_thread_pool_interface(std::shared_ptr<storage_thread_pool> thread_pool, name_set must_have_all_names = {},
                                  name_set must_have_any_names = {}, name_set good_to_have_names = {})
        : _pool(std::move(thread_pool)),
          _must_all(std::move(must_have_all_names)),
          _must_any(std::move(must_have_any_names)),
          _good(std::move(good_to_have_names)) {}

    // good_to_have_names are added to the interface's own
    template <class F, class... Args>
    std::future<std::invoke_result_t<F, storage&, Args...>> enqueue(F&& f, name_set good_to_have_names,
                                                                   Args&&... args) {
        good_to_have_names.insert(_good.begin(), _good.end());
        return _pool->enqueue(std::forward<F>(f), _must_all, _must_any, std::move(good_to_have_names),
                              std::forward<Args>(args)...);
    }
    void clear_from_storages(const name_set& names_to_clear) {
        _pool->clear_from_storages(_must_all, _must_any, names_to_clear);
    }
    // worker count of the underlying pool
    size_t nthreads() const { return _pool->nthreads(); }
    storage_thread_pool& pool() { return *_pool; }

private:
    std::shared_ptr<storage_thread_pool> _pool;
    name_set _must_all, _must_any, _good;
};

}  // namespace threading
}  // namespace hasty
