defmodule NxStereo.HipSegment.Server do
  @moduledoc """
  The process that owns a libnxs_hip context (one HIP device's streams and memory) for the
  `NxStereo.HipSegment`s that run through it, in the place `NxStereo.Matcher` holds the
  stereo_match Port (lib/nx_stereo/matcher.ex): a segment is configured on its first pair
  and again once its version changes; it is released when the process that runs it exits;
  a pair's reply goes to the caller of `send_run/6` as a `:gen_server` response. Unlike the
  Port, the context computes the pairs of different segments at once on the device, and
  answers each as it completes; a segment's own pairs stay in order.

  Options: `:name` (`#{inspect(__MODULE__)}`), `:device` (0).
  """
  use GenServer

  alias NxStereo.HipSegment.{Nif, Records}

  def start_link(opts \\ []),
    do: GenServer.start_link(__MODULE__, opts, name: Keyword.get(opts, :name, __MODULE__))

  @doc """
  Sends `request` (a run record) for segment `segment_id`, configured first with
  `configuration` unless the context has it at `version`, without waiting. The reply:
  `{:ok, reply_record}`, `{:error, {:configuration_refused, reason}}` or
  `{:error, {:refused, reason}}`.
  """
  def send_run(server, owner, segment_id, version, configuration, request),
    do: :gen_server.send_request(server, {:run, owner, segment_id, version, configuration, request})

  @impl GenServer
  def init(opts) do
    case Nif.open(Keyword.get(opts, :device, 0)) do
      {:ok, context} ->
        {:ok, %{context: context, configured: %{}, refused: %{}, owners: %{}, waiting: %{}}}

      {:error, reason} ->
        {:stop, {:no_device, reason}}
    end
  end

  @impl GenServer
  def handle_call({:run, owner, id, version, configuration, request}, from, state) do
    state = watch(state, owner, id)

    state =
      if Map.get(state.configured, id) == version do
        state
      else
        {:ok, ref} = Nif.submit(state.context, configuration)
        %{state | configured: Map.put(state.configured, id, version),
                  waiting: Map.put(state.waiting, ref, {:configure, id})}
      end

    case Nif.submit(state.context, request) do
      {:ok, ref} -> {:noreply, %{state | waiting: Map.put(state.waiting, ref, {:run, id, from})}}
      {:error, reason} -> {:reply, {:error, {:refused, reason}}, state}
    end
  end

  @impl GenServer
  def handle_info({:nxs_reply, ref, reply}, state) do
    {waiting, rest} = Map.pop(state.waiting, ref)
    {:noreply, answer(waiting, reply, %{state | waiting: rest})}
  end

  def handle_info({:DOWN, _ref, :process, owner, _reason}, state) do
    {ids, owners} = Map.pop(state.owners, owner, MapSet.new())
    for id <- ids, Map.has_key?(state.configured, id), do: Nif.submit(state.context, Records.release(id))
    {:noreply, %{state | owners: owners, configured: Map.drop(state.configured, Enum.to_list(ids)),
                         refused: Map.drop(state.refused, Enum.to_list(ids))}}
  end

  def handle_info(_message, state), do: {:noreply, state}

  @impl GenServer
  def terminate(_reason, state), do: Nif.close(state.context)

  # A segment's configure reply comes before its run's (the context keeps a segment's order).
  defp answer({:configure, id}, reply, state) do
    case Records.decode(reply) do
      :configured -> %{state | refused: Map.delete(state.refused, id)}
      {:refused, reason} -> %{state | configured: Map.delete(state.configured, id),
                                      refused: Map.put(state.refused, id, reason)}
      _other -> state
    end
  end

  defp answer({:run, id, from}, reply, state) do
    result =
      case {Map.get(state.refused, id), reply} do
        {reason, _reply} when is_binary(reason) -> {:error, {:configuration_refused, reason}}
        {nil, <<9, reason::binary>>} -> {:error, {:refused, reason}}
        {nil, reply} -> {:ok, reply}
      end

    GenServer.reply(from, result)
    state
  end

  defp answer(nil, _reply, state), do: state

  defp watch(state, owner, id) do
    case state.owners do
      %{^owner => ids} -> put_in(state.owners[owner], MapSet.put(ids, id))
      _other ->
        Process.monitor(owner)
        put_in(state.owners[owner], MapSet.new([id]))
    end
  end
end
