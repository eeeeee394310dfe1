// Host-only checks of include/hasty/threading.hpp (the reference's
// threading.cppm:15-395): storage accessors and their errors, task routing by
// must_have_all / must_have_any / good_to_have, the optimality wait,
// exceptions through futures, clear_from_storages, and tasks enqueued during
// a clear.  Prints "ALL OK" and exits 0 when every check holds.
#include <hasty/threading.hpp>

#include <atomic>
#include <cstdio>
#include <string>

using namespace hasty::threading;
using clk = std::chrono::steady_clock;

static int failures = 0;
static void expect(bool ok, const char* what) {
    std::printf("[threading] %-60s %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) ++failures;
}
template <class F> static std::string error_of(F f) {
    try {
        f();
    } catch (const std::runtime_error& e) {
        return e.what();
    }
    return "";
}

// a storage that knows its own index and holds the given names
static storage named(int idx, std::initializer_list<const char*> names) {
    storage s;
    s.add<int>("idx", std::make_shared<int>(idx));
    for (auto n : names) s.add<int>(n, std::make_shared<int>(0));
    return s;
}

int main() {
    // ---- storage ----
    {
        storage s;
        s.add<int>("a", std::make_shared<int>(3));
        s.add<int>("a", std::make_shared<int>(4));  // kept: insert does not replace
        expect(s.get_ref_throw<int>("a") == 3, "add keeps an existing entry");
        expect(!s.get_ref<double>("a").has_value() && s.get_ptr<double>("a") == nullptr,
               "get_ref / get_ptr of another type give nothing");
        expect(s.get_ref<int>("a")->get() == 3 && *s.get_ptr<int>("a") == 3, "get_ref / get_ptr");
        expect(error_of([&] { s.get_ref_throw<int>("b"); }) == "Storage does not contain key: b",
               "get_ref_throw of a missing name");
        expect(error_of([&] { s.get_ptr_throw<double>("a"); }).rfind("Storage does not contain key: a with type: ", 0) == 0,
               "get_ptr_throw of another type");
        expect(error_of([&] { s.clear<double>("a"); }) != "" && s.exist("a"), "clear<T> of another type throws");
        expect(error_of([&] { s.clear_throw<int>("zz"); }) == "Storage does not contain key: zz",
               "clear_throw of a missing name");
        s.clear<int>("missing");  // no-op
        s.add<double>("b", std::make_shared<double>(1.0));
        expect(s.names() == name_set{"a", "b"}, "names");
        s.clear("a");
        expect(!s.exist("a") && s.exist("b"), "clear by name");
    }

    // ---- routing: must_have_all / must_have_any, values through futures ----
    {
        std::vector<storage> st;
        st.push_back(named(0, {"gpu", "plan"}));
        st.push_back(named(1, {"gpu"}));
        st.push_back(named(2, {}));
        storage_thread_pool pool(std::move(st), 200);
        auto who = [](storage& s, int add) { return s.get_ref_throw<int>("idx") + add; };
        bool all_ok = true, any_ok = true;
        for (int r = 0; r < 20; ++r) {
            all_ok &= pool.enqueue(who, {"gpu", "plan"}, {}, {}, 100).get() == 100;
            const int w = pool.enqueue(who, {}, {"plan", "gpu"}, {}, 0).get();
            any_ok &= (w == 0 || w == 1);
        }
        expect(all_ok, "must_have_all routes to the one storage with every name");
        expect(any_ok, "must_have_any routes to storages with one of the names");
        // good_to_have: with a long optimality wait, the task goes to the
        // storage that has the name even though the others are idle
        bool good_ok = true;
        for (int r = 0; r < 10; ++r) good_ok &= pool.enqueue(who, {}, {}, {"plan"}, 0).get() == 0;
        expect(good_ok, "good_to_have prefers the storage holding the name");
        // a good_to_have name nobody holds: runs anyway after the wait
        const auto t0 = clk::now();
        const int w = pool.enqueue(who, {}, {}, {"nobody"}, 0).get();
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t0).count();
        expect(w >= 0 && w <= 2 && ms >= 190, "unmatched good_to_have runs after the optimality wait");
        // no good_to_have names: starts at once
        const auto t1 = clk::now();
        (void)pool.enqueue(who, {}, {}, {}, 0).get();
        const auto ms1 = std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t1).count();
        expect(ms1 < 150, "a task without good_to_have names starts at once");
        // exceptions reach the caller
        auto f = pool.enqueue([](storage& s) { s.get_ref_throw<int>("absent"); }, {}, {}, {});
        expect(error_of([&] { f.get(); }) == "Storage does not contain key: absent", "exceptions go to the future");
        // many tasks, all run exactly once
        std::atomic<int> count{0};
        std::vector<std::future<void>> fs;
        for (int i = 0; i < 500; ++i) fs.push_back(pool.enqueue([&count](storage&) { ++count; }, {}, {}, {}));
        for (auto& x : fs) x.get();
        expect(count == 500, "500 tasks each run once");
    }

    // ---- clear_from_storages ----
    {
        std::vector<storage> st;
        st.push_back(named(0, {"dev", "plan", "img"}));
        st.push_back(named(1, {"dev", "plan"}));
        st.push_back(named(2, {"plan"}));
        auto pool = std::make_shared<storage_thread_pool>(std::move(st), 10);
        // a running task holds the clear until it finishes
        std::atomic<bool> finished{false};
        auto slow = pool->enqueue(
            [&finished](storage&) {
                std::this_thread::sleep_for(std::chrono::milliseconds(100));
                finished = true;
            },
            {}, {}, {});
        std::this_thread::sleep_for(std::chrono::milliseconds(10));
        pool->clear_from_storages({"dev"}, {}, {"plan"});
        expect(finished.load(), "clear waits for running tasks");
        const auto& ss = pool->storages();
        expect(!ss[0].exist("plan") && !ss[1].exist("plan") && ss[2].exist("plan") && ss[0].exist("img"),
               "clear removes the names only from matching storages");
        pool->clear_from_storages({}, {"img", "nothing"}, {"img"});
        expect(!ss[0].exist("img") && ss[0].exist("dev"), "clear with must_have_any");
        // a task enqueued during a clear starts only after it
        std::atomic<bool> cleared{false};
        std::atomic<bool> saw_cleared{false};
        auto long_task = pool->enqueue(
            [](storage&) { std::this_thread::sleep_for(std::chrono::milliseconds(80)); }, {}, {}, {});
        std::this_thread::sleep_for(std::chrono::milliseconds(10));
        std::thread clearer([&] {
            pool->clear_from_storages({}, {}, {"dev"});
            cleared = true;
        });
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        auto late = pool->enqueue([&](storage& s) { saw_cleared = !s.exist("dev"); }, {}, {}, {});
        late.get();
        clearer.join();
        long_task.get();
        expect(saw_cleared.load() && cleared.load(), "a task enqueued during a clear runs after it");
        // the interface adds its requirement sets
        storage_thread_pool_interface iface(pool, {"plan"}, {}, {});
        const int w = iface.enqueue([](storage& s) { return s.get_ref_throw<int>("idx"); }, {}).get();
        expect(w == 2 && iface.nthreads() == 3, "interface routes by its must_have_all names");
        iface.clear_from_storages({"plan"});
        expect(!ss[2].exist("plan"), "interface clear uses its match sets");
    }

    // ---- shutdown with queued work: futures report broken promises ----
    {
        std::vector<storage> st;
        st.push_back(named(0, {}));
        std::future<int> never;
        {
            storage_thread_pool pool(std::move(st), 10);
            never = pool.enqueue([](storage&) { return 1; }, {"nonexistent"}, {}, {});
        }
        bool broken = false;
        try {
            never.get();
        } catch (const std::future_error&) {
            broken = true;
        }
        expect(broken, "a task no storage can take is abandoned at shutdown");
    }

    std::printf("[threading] %s (%d failures)\n", failures ? "FAILED" : "ALL OK", failures);
    return failures ? 1 : 0;
}
