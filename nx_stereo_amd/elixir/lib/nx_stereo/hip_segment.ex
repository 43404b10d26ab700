defmodule NxStereo.HipSegment do
  @moduledoc """
  A list of stages computed on an AMD GPU (MI355X, gfx950) by libnxs_hip, run on each
  pair as one request: `NxStereo.NativeSegment`'s counterpart for the GPU
  (lib/nx_stereo/native_segment.ex:1-246), with the same records, settings and outputs,
  through `NxStereo.HipSegment.Server` and its NIF instead of the stereo_match Port.

  Every stage NativeSegment accepts is accepted, with the same results bit for bit (f32
  values computed in the same order); the Disparity stage also takes 192 and 256
  disparities and search coarse_to_fine. The context keeps what the stages carry from one
  pair to the next on the device. Building a segment does not reach the device: it is
  configured on its first pair, and a configuration the library refuses answers that pair
  with `{:stop, reason}`. Pairs sent with `send_request/2` are computed in the order sent;
  pairs of other segments run beside them on the device.
  """

  @behaviour NxStereo.Segment

  alias NxStereo.{Segment, Stage}
  alias NxStereo.HipSegment.{Records, Server}

  @enforce_keys [:server, :id, :version, :members, :specs, :configuration, :inputs, :geometry]
  defstruct [:server, :owner, :id, :version, :members, :specs, :configuration, :inputs, :geometry,
             times_us: []]

  @type t :: %__MODULE__{}

  @doc "Where a HIP segment runs: an AMD GPU, through libnxs_hip on ROCm."
  @impl Segment
  def compute(%__MODULE__{}), do: %{runs_on: :gpu, implementation: :hip, backend: :rocm}

  @doc "Three pairs of a segment in flight at once: its run slots in the context."
  @impl Segment
  def depth, do: 3

  @impl Segment
  def accepts?(module, state, compute),
    do: compute == :native and refusal(%{module: module, state: state}) == nil

  @doc "The segment of `members` for `geometry`; `opts`: `:server` (`NxStereo.HipSegment.Server`)."
  @impl Segment
  def build(members, geometry, opts) do
    with {:ok, opts} <- Keyword.validate(opts, server: Server),
         nil <- Enum.find_value(members, &refusal/1),
         {:ok, specs} <- Segment.entry_specs(members, geometry) do
      segment = %__MODULE__{
        server: opts[:server],
        owner: self(),
        id: rem(System.unique_integer([:positive]), 0x1_0000_0000),
        version: 0,
        members: members,
        specs: specs,
        configuration: nil,
        inputs: nil,
        geometry: geometry
      }

      {:ok, configure(segment)}
    else
      {:error, unknown} when is_list(unknown) -> {:error, "unknown build options #{inspect(unknown)}"}
      error -> error
    end
  end

  defp refusal(%{module: module, state: state}) do
    cond do
      Stage.native_kind(module) == nil -> {:error, "#{inspect(module)} has no native implementation"}
      Stage.compute(module, state) == :nx -> {:error, "#{inspect(module)} is computed in Nx, as its settings ask"}
      true -> nil
    end
  end

  @impl Segment
  def reload(%__MODULE__{} = segment, members, geometry),
    do: configure(%{segment | members: members, geometry: geometry, version: segment.version + 1})

  @impl Segment
  def put_member_state(%__MODULE__{members: members} = segment, index, state) do
    members = List.update_at(members, index, &%{&1 | state: state})
    %{segment | members: members, inputs: Records.inputs(members, segment.geometry)}
  end

  @doc "`current`: the context keeps what the stages carry, on the device."
  @impl Segment
  def merge_carry(%__MODULE__{} = current, %__MODULE__{}), do: current

  @impl Segment
  def send_request(entries, %__MODULE__{} = segment) do
    with {:ok, entry} <- Segment.entry_tensors(entries, segment.specs) do
      request = Records.run(segment.id, segment.inputs, entry)

      {:ok, Server.send_run(segment.server, segment.owner, segment.id, segment.version,
                            segment.configuration, request)}
    end
  end

  @impl Segment
  def response(message, request_id, %__MODULE__{} = segment) do
    case :gen_server.check_response(message, request_id) do
      :no_reply -> :no_reply
      response -> received(response, segment)
    end
  end

  @impl Segment
  def await(request_id, %__MODULE__{} = segment, timeout) do
    case :gen_server.receive_response(request_id, timeout) do
      :timeout -> :timeout
      response -> received(response, segment)
    end
  end

  defp received({:reply, {:error, {:configuration_refused, reason}}}, _segment), do: {:stop, reason}
  defp received({:reply, {:error, _reason} = error}, _segment), do: error

  defp received({:reply, {:ok, reply}}, segment) do
    case Records.decode(reply) do
      {:outputs, times_us, outputs} ->
        {:ok, outputs, %{segment | times_us: times_us}, System.monotonic_time(:nanosecond)}

      {:refused, reason} -> {:error, reason}
      {:error, _reason} = error -> error
      _other -> {:error, :bad_reply}
    end
  end

  defp received({:error, {reason, _server}}, _segment), do: {:error, {:server_down, reason}}

  @doc "Each member's name and GPU time for the last pair, in nanoseconds."
  @impl Segment
  def member_times(%__MODULE__{members: members, times_us: times_us}),
    do: Enum.zip_with(members, times_us, &{&1.name, &2 * 1_000})

  @doc "`outputs`, which the reply brought back."
  @impl Segment
  def collect(outputs, %__MODULE__{}, _timeout), do: {:ok, outputs}

  defp configure(segment) do
    %{segment | configuration: Records.configure(segment.id, segment.members, segment.geometry),
                inputs: Records.inputs(segment.members, segment.geometry)}
  end
end
