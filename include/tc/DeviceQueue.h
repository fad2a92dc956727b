// Drop-in of TC/DeviceQueue.h:6-16 over a tc_stream.
#pragma once
#include "platform.h"

class DeviceQueue {
  public:
    explicit DeviceQueue(QueueInfo info) : queue(info.cmd_queue) {}
    tDeviceRC finish() { return tc_stream_sync(queue); }   // clFinish, TC/DeviceQueue.cpp:14-24
    tDeviceRC flush() { return tc_stream_flush(queue); }   // clFlush, TC/DeviceQueue.cpp:26-35
  private:
    tc_stream queue;
};
