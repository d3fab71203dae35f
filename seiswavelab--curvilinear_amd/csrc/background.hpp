// background.hpp — the worker thread behind the background-stepping façades (async2d.cpp, async3d.cpp), with the
// semantics of the reference's SeisWave::BackgroundRunner (include/SeisCommon/BackgroundRunner.hpp): one
// persistent thread, one task at a time, cancellation between steps, and only the owner consumes the finished
// result (a failure is rethrown to the owner when it consumes).
#pragma once
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace swl {

// One persistent worker; only the owner consumes finished results; cancellation between steps.
template <class Publication> class BackgroundWorker {
    std::mutex mutex_;
    std::condition_variable wake_, done_;
    std::function<Publication(const std::atomic_bool &)> task_;
    std::optional<Publication> result_;
    std::string failure_;
    std::atomic_bool cancel_{false};
    bool busy_ = false, ready_ = false, shutdown_ = false;
    std::thread thread_;

    std::optional<Publication> consume() { // mutex held
        if (!ready_) return {};
        ready_ = busy_ = false;
        if (!failure_.empty()) throw std::runtime_error(std::exchange(failure_, {}));
        auto r = std::move(result_);
        result_.reset();
        return r;
    }

public:
    BackgroundWorker()
        : thread_([this] {
              for (;;) {
                  std::unique_lock<std::mutex> lock(mutex_);
                  wake_.wait(lock, [&] { return shutdown_ || static_cast<bool>(task_); });
                  if (shutdown_) return;
                  auto task = std::move(task_);
                  task_ = nullptr;
                  lock.unlock();
                  std::optional<Publication> value;
                  std::string failure;
                  try {
                      value = task(cancel_);
                  } catch (const std::exception &e) {
                      failure = e.what();
                  }
                  lock.lock();
                  result_ = std::move(value), failure_ = failure, ready_ = true;
                  done_.notify_all();
              }
          }) {}
    ~BackgroundWorker() {
        cancel();
        {
            std::unique_lock<std::mutex> lock(mutex_);
            done_.wait(lock, [&] { return !busy_ || ready_; });
            shutdown_ = true;
            wake_.notify_all();
        }
        thread_.join();
    }
    bool busy() {
        std::lock_guard<std::mutex> lock(mutex_);
        return busy_;
    }
    bool submit(std::function<Publication(const std::atomic_bool &)> task) {
        std::lock_guard<std::mutex> lock(mutex_);
        if (busy_) return false;
        cancel_ = false, busy_ = true, task_ = std::move(task);
        wake_.notify_one();
        return true;
    }
    void cancel() { cancel_ = true; }
    std::optional<Publication> poll() {
        std::lock_guard<std::mutex> lock(mutex_);
        return consume();
    }
    std::optional<Publication> wait() {
        cancel();
        std::unique_lock<std::mutex> lock(mutex_);
        done_.wait(lock, [&] { return !busy_ || ready_; });
        return consume();
    }
};

} // namespace swl
